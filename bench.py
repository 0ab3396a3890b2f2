"""Benchmark: bubble-substeps/sec of the world step on one MI355X.

Workload (default "batched"): 4,096 independent worlds with bubble counts log-uniform in
[16, 4096] (about 3.0M bubbles), radii U[6, 20], 0-16 kinematic bodies each, walls, advanced
together: the timed call advances every world by `--steps` frames of dt = 1/60, one launch per
kernel class, each workgroup carrying its world through all the frames. Beside it: the plain
single-world step (1,024 bubbles + 64 kinematic bodies, one launch per frame), the 262,144-bubble
world of config 3 on the cooperative grid, and the CPU oracle on the host cores.

One JSON line on stdout. value = bubble-substeps/s = sum over worlds and frames of
(bubbles x substeps) / wall time of the timed call, with every world resident on the GPU first.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Hardware queues of this process, read by the HIP runtime when it starts (before any device
# call below). The batched step runs each kernel class on its own stream and the server keeps one
# launch on each of its 32 streams; with HIP's default of 4 queues those streams would share
# queues and their launches would run one after another.
HW_QUEUES = 32
os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)

HBM_PEAK_GBPS = 8000.0


def bytes_per_bubble_substep(mean_degree: float) -> float:
    """Compulsory HBM bytes per bubble-substep: 168 + 8 * mean adjacency-row length (positions,
    velocities, accelerations, radius, stiffness, inverse mass, flags, margin, x0 read and written
    across the kick-drift, force and reduction passes, plus the adjacency row read twice)."""
    return 168.0 + 8.0 * mean_degree


def kernel_name(k, res):
    """Kernel of class k: the LDS configurations, the single-workgroup global-state kernel (class 10)
    and the cooperative kernel (class 11)."""
    if k == 10:
        return "world_step_kernel<1024> (global state, one workgroup)"
    if k == 11:
        return "world_coop_kernel<256> (cooperative grid)"
    return f"world_lds_kernel<{int(res['class_threads'][k])} threads, {int(res['class_capacity'][k])} bubbles>"


def percentile(values, q):
    return float(np.percentile(np.asarray(values, np.float64), q)) if len(values) else float("nan")


def build_batched(worlds: int):
    from bubblephysics_amd import scenes
    return scenes.config4(worlds=worlds)


def run_batched(args, scenes_list):
    """Throughput: one call advances every world `steps` steps (one launch per kernel class, each
    workgroup carrying its world through all the steps). Latency: `latency_steps` calls of one step."""
    from bubblephysics_amd import solver
    batch = solver.Batch(scenes_list)
    sizes = np.array([len(s.bubbles) for s in scenes_list], np.int64)
    if args.warmup:
        batch.step(args.warmup)
    classes = int(batch._lib.bp_batch_class_count())
    t0 = time.perf_counter()
    infos = batch.step(args.steps)
    elapsed = time.perf_counter() - t0
    ms, _, threads, capacity = batch.last_class_timing()
    ok = infos["status"] == 0                      # failed and unexecuted steps do not count
    substeps = np.where(ok, infos["substep_count"], 0).astype(np.int64)
    substep_bubbles = int((substeps * sizes[None, :]).sum())
    deg = infos["adjacency_entries"].astype(np.float64) / np.maximum(sizes[None, :], 1)
    per = substeps * sizes[None, :] * bytes_per_bubble_substep(deg)
    class_bytes = np.zeros(classes)
    for k in range(classes):
        class_bytes[k] = float(per[:, infos[0]["kernel_class"] == k].sum())
    failed = int((infos[-1]["status"] != 0).sum())
    latencies = []
    for _ in range(args.latency_steps):
        s0 = time.perf_counter()
        batch.step(1)
        latencies.append((time.perf_counter() - s0) * 1e3)
    return {
        "class_threads": threads, "class_capacity": capacity, "elapsed": elapsed, "substep_bubbles": substep_bubbles,
        "latencies": latencies, "class_ms": ms, "class_bytes": class_bytes, "failed_worlds": failed,
        "bubbles": int(sizes.sum()), "worlds": len(sizes), "overflow": int(infos["candidate_overflow"].max()),
    }


def run_large(steps: int):
    """The 262,144-bubble world of config 3 on the cooperative grid."""
    from bubblephysics_amd import scenes, solver
    scene = scenes.config3()
    world = solver.NativeWorld(scene)
    n = len(scene.bubbles)
    world.step(0)
    total, statuses, gpu_ms = 0, [], 0.0
    t0 = time.perf_counter()
    for k in range(1, steps + 1):
        f = world.step(k)
        statuses.append(f.status)
        total += n * f.substep_count
        gpu_ms += f.gpu_milliseconds
    elapsed = time.perf_counter() - t0
    return {"value": total / elapsed, "ms_per_step": elapsed / steps * 1e3, "gpu_ms_per_step": gpu_ms / steps,
            "bubbles": n, "bodies": len(scene.bodies), "statuses": sorted(set(statuses))}


def run_served(args):
    """Serving: Poisson step requests (10% with transactions) over resident worlds, continuous batching."""
    from bubblephysics_amd import serving, solver
    t = time.perf_counter()
    worlds = serving.serving_worlds(args.serve_worlds, seed=20261206)
    print(f"[bench] built {len(worlds)} serving worlds in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    batch = solver.Batch(worlds)
    batch.step(1)                                          # packs and warms every kernel class
    server = serving.Server(batch, memory_budget=int(200e9), streams=args.serve_streams,
                            max_inflight_per_class=args.serve_inflight)
    requests = serving.poisson_requests(len(worlds), args.serve_rate, args.serve_seconds, seed=20261207)
    print(f"[bench] serving {len(requests)} requests", file=sys.stderr, flush=True)
    stats = server.run(requests).summary()
    server.close()
    stats["bubbles"] = int(sum(len(w.bubbles) for w in worlds))
    stats["worlds"] = len(worlds)
    return stats


def run_served_batch(args, scenes_list):
    """Serving the worlds of the batched workload: Poisson step requests (10% with transactions) at
    `serve_c4_rate` over `serve_c4_seconds`, from a fresh copy of the worlds."""
    from bubblephysics_amd import serving, solver
    batch = solver.Batch(scenes_list)
    batch.step(1)
    server = serving.Server(batch, streams=args.serve_streams, max_inflight_per_class=args.serve_inflight)
    requests = serving.poisson_requests(len(scenes_list), args.serve_c4_rate, args.serve_c4_seconds, seed=20261208)
    stats = server.run(requests).summary()
    server.close()
    batch.close()
    return stats


def run_plain(steps: int):
    from bubblephysics_amd import scenes, solver
    scene = scenes.config2()
    world = solver.NativeWorld(scene)
    n = len(scene.bubbles)
    for k in range(3):
        world.step(k)
    lat, total = [], 0
    t0 = time.perf_counter()
    for k in range(steps):
        s0 = time.perf_counter()
        f = world.step(k)
        lat.append((time.perf_counter() - s0) * 1e3)
        total += n * f.substep_count
    elapsed = time.perf_counter() - t0
    return {"value": total / elapsed, "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99), "bubbles": n}


def cpu_baseline(scenes_list, budget_s: float, threads: int, steps: int = 3):
    """The oracle (FP32) on the worlds of the same batch, `steps` steps each from their initial
    state, spread over `threads` host threads (the oracle library runs outside the interpreter
    lock), until the wall budget is used. Executed bubble-substeps of successful steps / wall time."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import bp_oracle
    bp_oracle.load()
    deadline = time.perf_counter() + budget_s

    def one(s):
        if time.perf_counter() >= deadline:
            return None
        _, _, frames, _ = bp_oracle.run(s, steps, assert_mode=1)
        ok = frames["status"] == 0
        return len(s.bubbles) * int(frames["substeps"][ok].sum())

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as pool:
        done = [r for r in pool.map(one, scenes_list) if r is not None]
    used = time.perf_counter() - t0
    return {"value": sum(done) / used if used > 0 else float("nan"), "worlds": len(done), "seconds": used,
            "threads": threads, "steps": steps}


def fp64_drift(checkpoints=(1, 10, 30, 60)):
    """Largest difference of bubble positions between the FP32 and FP64 oracle runs of the
    256-bubble world (config 1) after each checkpoint's number of steps, absolute and relative to
    the region size. The pile is chaotic, so the difference grows with the step count."""
    from bubblephysics_amd import scenes
    from oracle import bp_oracle
    scene = scenes.config1()
    r = scene.config.region
    size = float(max(r[2] - r[0], r[3] - r[1]))
    out = {"world": "config 1: 256 bubbles, walls, gravity"}
    for steps in checkpoints:
        b32, _, f32, _ = bp_oracle.run(scene, steps, precision=32, assert_mode=1)
        b64, _, f64, _ = bp_oracle.run(scene, steps, precision=64, assert_mode=1)
        d = np.abs(b32["center"].astype(np.float64) - b64["center"].astype(np.float64)).max(axis=1)
        out[f"after_{steps}_steps"] = {"max_abs": float(d.max()), "median_abs": float(np.median(d)),
                                       "max_rel_to_region": float(d.max()) / size, "completed": [len(f32), len(f64)]}
    return out


def copy_calibration(nbytes: int, repeats: int = 5):
    """A plain device copy of `nbytes`, for scaling the FETCH_SIZE and WRITE_SIZE counters of a
    profiler run: the counters of this kernel should read nbytes / 1024 KiB each."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        dst.copy_(src)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / repeats * 1e3
    print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": ms, "GB/s": 2 * nbytes / ms / 1e6}))


def recorded_traffic(kernel: str, steps: int, worlds: int):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by the copy calibration) as
    recorded in profiles/traffic.txt for this workload, or None when no record matches."""
    path = ROOT / "profiles" / "traffic.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split("|")
        if len(parts) != 5 or line.lstrip().startswith("#"):
            continue
        name, st, wd, fetch, write = (p.strip() for p in parts)
        if name == kernel and st == str(steps) and wd == str(worlds):
            return float(fetch) + float(write)
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--plain-steps", type=int, default=60)
    ap.add_argument("--latency-steps", type=int, default=10)
    ap.add_argument("--large-steps", type=int, default=5)
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--serve-rate", type=float, default=20000.0)
    ap.add_argument("--serve-worlds", type=int, default=8192)
    ap.add_argument("--serve-c4-seconds", type=float, default=5.0)
    ap.add_argument("--serve-c4-rate", type=float, default=20000.0)
    ap.add_argument("--serve-streams", type=int, default=32)
    ap.add_argument("--serve-inflight", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=min(os.cpu_count() or 1, 16))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=0, help="only run the copy calibration kernel of this size")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the worlds are independent, so a multi-GPU run would shard them by "
              f"world with no exchange; only one MI355X is measured (--gpus {args.gpus} is not supported)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_bytes:
        copy_calibration(args.copy_bytes)
        return

    t_build = time.perf_counter()
    scenes_list = build_batched(args.worlds)
    print(f"[bench] built {len(scenes_list)} worlds in {time.perf_counter() - t_build:.1f}s", file=sys.stderr, flush=True)
    res = run_batched(args, scenes_list)
    print(f"[bench] batched done: {res['elapsed']:.2f}s for {args.steps} steps", file=sys.stderr, flush=True)
    plain = run_plain(args.plain_steps) if args.plain_steps > 0 else {"value": float("nan"), "p50_ms": float("nan"),
                                                                        "p99_ms": float("nan"), "bubbles": 0}
    print(f"[bench] plain done: {plain['value']:.3g} bubble-substeps/s", file=sys.stderr, flush=True)
    large = run_large(args.large_steps) if args.large_steps > 0 else None
    if large:
        print(f"[bench] large done: {large['value']:.3g} bubble-substeps/s", file=sys.stderr, flush=True)
    served_c4 = run_served_batch(args, scenes_list) if args.serve_c4_seconds > 0 else None
    if served_c4:
        print(f"[bench] served batch worlds: {served_c4}", file=sys.stderr, flush=True)
    served = run_served(args) if args.serve_seconds > 0 else None
    if served:
        print(f"[bench] served: {served}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scenes_list, args.cpu_seconds, args.cpu_threads)
        cpu["drift"] = fp64_drift()
        print(f"[bench] cpu done: {cpu['value']:.3g} bubble-substeps/s on {cpu['threads']} threads", file=sys.stderr,
              flush=True)

    value = res["substep_bubbles"] / res["elapsed"]
    k = int(np.argmax(res["class_ms"]))
    dur_s = float(res["class_ms"][k]) / 1e3          # one launch of the class covers all the timed steps
    achieved = res["class_bytes"][k] / dur_s / 1e9 if dur_s > 0 else 0.0
    traffic = recorded_traffic(kernel_name(k, res), args.steps, res["worlds"])
    line = {
        "metric": "bubble-substeps/sec on 1 MI355X (batched/served vs plain) + p50/p99 step latency",
        "value": value,
        "unit": "bubble-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched: {res['worlds']} worlds, bubbles log-uniform in [16, 4096] ({res['bubbles']} total), "
                        f"radii U[6,20], 0-16 kinematic bodies per world, walls, dt 1/60",
            "worlds": res["worlds"], "bubbles": res["bubbles"], "parallelism": "one workgroup per world",
        },
        "p50_step_ms": percentile(res["latencies"], 50),     # one step of every world, host clock
        "p99_step_ms": percentile(res["latencies"], 99),
        "failed_worlds": res["failed_worlds"],
        "candidate_overflow_max": res["overflow"],
        "plain": {"workload": "plain: 1 world, 1024 bubbles + 64 kinematic bodies", "value": plain["value"],
                  "p50_step_ms": plain["p50_ms"], "p99_step_ms": plain["p99_ms"]},
        "batched_over_plain": value / plain["value"] if plain["value"] > 0 else None,   # NaN compares false
        "large": None if large is None else {
            "workload": f"config 3: 1 world, {large['bubbles']} bubbles + {large['bodies']} bodies (kinematic and walls), "
                        "cooperative grid", "value": large["value"], "ms_per_step": large["ms_per_step"],
            "gpu_ms_per_step": large["gpu_ms_per_step"], "statuses": large["statuses"]},
        "served_batch_worlds": None if served_c4 is None else {
            "workload": f"serving the {res['worlds']} batched worlds above: Poisson {args.serve_c4_rate:.0f} step "
                        f"requests/s for {args.serve_c4_seconds:.0f}s, 10% with transactions",
            "over_plain": served_c4["bubble_substeps_per_s"] / plain["value"] if plain["value"] > 0 else None,
            **served_c4},
        "served": None if served is None else {
            "workload": f"serving: {served['worlds']} resident worlds, bubbles log-uniform in [32, 16384] "
                        f"({served['bubbles']} total), Poisson {args.serve_rate:.0f} step requests/s for "
                        f"{args.serve_seconds:.0f}s, 10% with transactions, budget 200 GB",
            **{k: v for k, v in served.items() if k not in ("worlds", "bubbles")}},
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
            "kernel": kernel_name(k, res),
            "kernel_ms_per_launch": float(res["class_ms"][k]),
        },
        "cpu_baseline": None if cpu is None else {
            "value": cpu["value"], "unit": "bubble-substeps/s", "cores": cpu["threads"], "kind": "port",
            "sample": f"oracle FP32, {cpu['steps']} steps of each of the first {cpu['worlds']} worlds of the same "
                      f"batch from their initial state, worlds spread over {cpu['threads']} host threads "
                      f"({cpu['seconds']:.1f}s wall, about {cpu['seconds'] * cpu['threads']:.0f} thread-seconds)",
            "fp64_drift": cpu["drift"],
        },
    }
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
