"""Benchmark: bubble-substeps/sec of the world step on one MI355X.

Workload of `value` ("batched"): 4,096 independent worlds with bubble counts log-uniform in
[16, 4096] (about 3.0M bubbles), radii U[6, 20], 0-16 kinematic bodies each, walls, advanced
together: the timed call replays hipGraphs that advance every world by `--steps` frames of
dt = 1/60, captured after `--warmup` untimed frames: one launch per kernel class (workgroup pools
for the classes whose workgroups fill a CU), each workgroup carrying its world through all the
frames. Beside it, on the same box:
  * the same launches made directly, without pools, and after the autotuner;
  * the plain single-world step (one launch per frame) of the 256-bubble world and of the
    1,024-bubble + 64-body world;
  * the 262,144-bubble world on the cooperative grid;
  * serving the batched worlds at rising offered loads, and the 8,192-world serving workload
    (60 s), then again with a device-memory budget that forces evictions;
  * the CPU oracle on the host cores: the batch, and the single worlds on one thread and with its
    per-phase threads;
  * the rigid solver for worlds without bubbles: 4,096 seeded rigid crowds in one launch
    (body-steps/s), with the oracle's rigid solver on the host cores beside it.

One JSON line on stdout. value = bubble-substeps/s = sum over worlds and frames of
(bubbles x substeps) / wall time of the timed call, with every world resident on the GPU first.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import platform
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
HOST_THREADS = 16          # host threads a one-GPU job may keep busy


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bytes_per_bubble_substep(mean_degree: float) -> float:
    """Compulsory HBM bytes per bubble-substep: 168 + 8 * mean adjacency-row length (positions,
    velocities, accelerations, radius, stiffness, inverse mass, flags, margin, x0 read and written
    across the kick-drift, force and reduction passes, plus the adjacency row read twice)."""
    return 168.0 + 8.0 * mean_degree


def kernel_name(k, threads, capacity):
    """Kernel of class k: the LDS configurations, the single-workgroup global-state kernel (class 10)
    and the cooperative kernel (class 11)."""
    if k == 10:
        return "world_step_kernel<1024> (global state, one workgroup)"
    if k == 11:
        return "world_coop_kernel<256> (cooperative grid)"
    return f"world_lds_kernel<{int(threads[k])} threads, {int(capacity[k])} bubbles>"


def percentile(values, q):
    return float(np.percentile(np.asarray(values, np.float64), q)) if len(values) else float("nan")


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# ---------------------------------------------------------------------------------------------
# GPU measurements
# ---------------------------------------------------------------------------------------------
def summarize(infos, sizes, elapsed):
    ok = infos["status"] == 0                      # failed and unexecuted steps do not count
    substeps = np.where(ok, infos["substep_count"], 0).astype(np.int64)
    return int((substeps * sizes[None, :]).sum()), substeps


def run_batched(args, scenes_list, *, mode="graph", label="batched"):
    """Throughput: every world advances `steps` steps in one timed call. mode "graph": the steps
    captured into hipGraphs after the warm-up (one graph per kernel class, each replayed on its
    class's stream, pools planned from the warm-up's measured times) and replayed once; "direct":
    the same launches made directly; "unpooled": direct launches of one workgroup per world.
    Latency: `latency_steps` direct calls of one step."""
    from bubblephysics_amd import solver
    batch = solver.Batch(scenes_list)
    batch.set_pooling(mode != "unpooled")
    sizes = np.array([len(s.bubbles) for s in scenes_list], np.int64)
    if args.warmup:
        batch.step(args.warmup)
    classes = int(batch._lib.bp_batch_class_count())
    if mode == "graph":
        batch.capture(args.steps)
        t0 = time.perf_counter()
        infos = batch.replay()
        elapsed = time.perf_counter() - t0
    else:
        t0 = time.perf_counter()
        infos = batch.step(args.steps)
        elapsed = time.perf_counter() - t0
    ms, _, threads, capacity = batch.last_class_timing()
    work, substeps = summarize(infos, sizes, elapsed)
    deg = infos["adjacency_entries"].astype(np.float64) / np.maximum(sizes[None, :], 1)
    per = substeps * sizes[None, :] * bytes_per_bubble_substep(deg)
    class_bytes = np.zeros(classes)
    for k in range(classes):
        class_bytes[k] = float(per[:, infos[0]["kernel_class"] == k].sum())
    latencies = []
    for _ in range(args.latency_steps):
        s0 = time.perf_counter()
        batch.step(1)
        latencies.append((time.perf_counter() - s0) * 1e3)
    batch.close()
    log(f"{label}: {elapsed * 1e3:.1f} ms for {args.steps} steps, {work / elapsed:.4g} bubble-substeps/s")
    return {
        "class_threads": threads, "class_capacity": capacity, "elapsed": elapsed, "substep_bubbles": work,
        "latencies": latencies, "class_ms": ms, "class_bytes": class_bytes, "failed_worlds": int((infos[-1]["status"] != 0).sum()),
        "bubbles": int(sizes.sum()), "worlds": len(sizes), "overflow": int(infos["candidate_overflow"].max()),
    }


def run_autotuned(args, scenes_list):
    """The autotuner's trials on the warmed-up batch (whole-batch steps per kernel configuration and
    bucket count), then the same timed steps with its choices."""
    from bubblephysics_amd import autotune, solver
    batch = solver.Batch(scenes_list)
    sizes = np.array([len(s.bubbles) for s in scenes_list], np.int64)
    if args.warmup:
        batch.step(args.warmup)
    t = time.perf_counter()
    table = autotune.tune(batch, repeats=args.tune_repeats)
    tune_s = time.perf_counter() - t
    batch.step(1)                                  # repacks with the choices and measures the worlds again
    t0 = time.perf_counter()
    infos = batch.step(args.steps)
    elapsed = time.perf_counter() - t0
    work, _ = summarize(infos, sizes, elapsed)
    batch.close()
    moved = [{"bubbles": [int(r["min_bubbles"]), int(r["max_bubbles"])], "class": int(r["best_class"]),
              "bucket_shift": int(r["best_bucket_shift"])} for r in table if r["worlds"]]
    log(f"autotuned: {elapsed * 1e3:.1f} ms, {work / elapsed:.4g} (trials {tune_s:.1f}s)")
    return {"value": work / elapsed, "ms_per_step": elapsed / args.steps * 1e3, "trials_s": tune_s, "choices": moved,
            "note": "timed one step later than the other lines (the step that adopts the choices)"}


def run_plain(scene, steps, warmup=3):
    """One world, one launch per frame, frames timed one by one on the host clock."""
    from bubblephysics_amd import solver
    world = solver.NativeWorld(scene)
    n = len(scene.bubbles)
    for k in range(warmup):
        world.step(k)
    lat, total, statuses = [], 0, set()
    t0 = time.perf_counter()
    for k in range(steps):
        s0 = time.perf_counter()
        f = world.step(warmup + k)
        lat.append((time.perf_counter() - s0) * 1e3)
        statuses.add(int(f.status))
        if f.status == 0:
            total += n * f.substep_count
        else:
            break
    elapsed = time.perf_counter() - t0
    return {"value": total / elapsed, "p50_step_ms": percentile(lat, 50), "p99_step_ms": percentile(lat, 99),
            "bubbles": n, "steps": len(lat), "statuses": sorted(statuses)}


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


def run_served_sweep(args, scenes_list, plain_value):
    """Serving the batched worlds at rising offered loads (Poisson step requests, 10% with
    transactions), each for `serve_c4_seconds`, one server over the same resident worlds. The
    highest throughput is the server's ceiling on this workload; the line with p99 <= 4 ms is the
    highest load the latency target allows."""
    from bubblephysics_amd import serving, solver
    batch = solver.Batch(scenes_list)
    batch.step(1)
    server = serving.Server(batch, streams=args.serve_streams, max_inflight_per_class=args.serve_inflight)
    out = []
    for j, rate in enumerate(args.serve_c4_rates):
        requests = serving.poisson_requests(len(scenes_list), rate, args.serve_c4_seconds, seed=20261208 + j)
        stats = server.run(requests).summary()
        stats["offered_per_s"] = rate
        stats["over_plain"] = stats["bubble_substeps_per_s"] / plain_value if plain_value > 0 else None
        out.append(stats)
        log(f"served batch worlds at {rate:.0f}/s: {stats['bubble_substeps_per_s']:.4g} bubble-substeps/s, "
            f"p50 {stats['p50_ms']:.2f} ms, p99 {stats['p99_ms']:.2f} ms")
    server.close()
    batch.close()
    best = max(out, key=lambda s: s["bubble_substeps_per_s"])
    within = [s for s in out if s["p99_ms"] <= 4.0]
    return {"sweep": out, "ceiling_bubble_substeps_per_s": best["bubble_substeps_per_s"],
            "ceiling_over_plain": best["over_plain"],
            "best_within_p99_4ms": max(within, key=lambda s: s["bubble_substeps_per_s"])["offered_per_s"] if within else None}


def run_served(args):
    """Config 5: Poisson step requests (10% with transactions) over 8,192 resident worlds for
    `serve_seconds`; then a shorter run with a device-memory budget below the resident footprint
    (`serve_evict_budget` of it), which forces admissions and LRU evictions to host memory."""
    from bubblephysics_amd import serving, solver
    t = time.perf_counter()
    worlds = serving.serving_worlds(args.serve_worlds, seed=20261206)
    log(f"built {len(worlds)} serving worlds in {time.perf_counter() - t:.1f}s")
    batch = solver.Batch(worlds)
    batch.step(1)                                          # packs and warms every kernel class
    server = serving.Server(batch, memory_budget=int(200e9), streams=args.serve_streams,
                            max_inflight_per_class=args.serve_inflight)
    requests = serving.poisson_requests(len(worlds), args.serve_rate, args.serve_seconds, seed=20261207)
    log(f"serving {len(requests)} requests")
    stats = server.run(requests).summary()
    server.close()
    resident = int(stats["resident_bytes"])
    stats["bubbles"] = int(sum(len(w.bubbles) for w in worlds))
    stats["worlds"] = len(worlds)
    log(f"served: {stats['bubble_substeps_per_s']:.4g} bubble-substeps/s, p99 {stats['p99_ms']:.2f} ms")
    evict = None
    if args.serve_evict_seconds > 0 and resident > 0:
        budget = int(resident * args.serve_evict_budget)
        server = serving.Server(batch, memory_budget=budget, streams=args.serve_streams,
                                max_inflight_per_class=args.serve_inflight)
        requests = serving.poisson_requests(len(worlds), args.serve_rate, args.serve_evict_seconds, seed=20261209)
        evict = server.run(requests).summary()
        server.close()
        evict["memory_budget_bytes"] = budget
        log(f"served under a {budget / 1e9:.1f} GB budget: {evict['evictions']} evictions, p99 {evict['p99_ms']:.2f} ms")
    batch.close()
    return stats, evict


# ---------------------------------------------------------------------------------------------
# CPU baselines (the oracle; test infrastructure used here only as the timed CPU restatement)
# ---------------------------------------------------------------------------------------------
def rigid_scenes(worlds: int, bodies: int):
    """Seeded rigid crowds (bubblephysics_amd.rigid.crowd): `bodies` dynamic bodies of four shapes,
    six statics and kinematics, five joints, gravity, in a walled 400 x 500 box."""
    from bubblephysics_amd import rigid
    out = []
    for w in range(worlds):
        objs, cfg, joints = rigid.crowd(50_000 + w, bodies=bodies)
        out.append(rigid.records_for(objs, cfg, joints=joints))
    return out


def run_rigid(args, scenes_rigid):
    """Rigid solver for worlds without bubbles: every world in one launch, all timed steps inside it.
    The untimed warm-up steps let the crowds fall and pile up, so the timed steps are contact-heavy.
    Returns the line and the warmed-up states of the first worlds (the CPU leg starts from them)."""
    from bubblephysics_amd import rigid
    batch = rigid.RigidBatch(scenes_rigid)
    batch.step(args.rigid_warmup)
    warmed = [dataclasses.replace(sc, bodies=batch.read_bodies(w)) for w, sc in enumerate(scenes_rigid[:256])]
    t0 = time.perf_counter()
    infos = batch.step(args.rigid_steps)
    elapsed = time.perf_counter() - t0
    bodies = sum(len(s.bodies) for s in scenes_rigid)
    ok = int((infos["status"] == 0).all(axis=0).sum())
    return {"value": bodies * args.rigid_steps / elapsed, "elapsed": elapsed, "gpu_ms": batch.last_gpu_ms(),
            "bodies": bodies, "worlds": len(scenes_rigid), "ok_worlds": ok,
            "pairs_per_step": float(infos["candidate_pairs"].sum(axis=1).mean()),
            "contacts": sum(len(batch.contacts(w)) for w in range(min(64, len(scenes_rigid)))) / min(64, len(scenes_rigid))}, warmed


def cpu_rigid(warmed, threads, steps, budget_s):
    """The oracle's rigid solver on the host cores, from the GPU line's warmed-up states (the same
    bits as the oracle's own warm-up): worlds spread over `threads` threads, each run for the GPU
    line's timed steps; as many worlds as fit about `budget_s` of work."""
    from oracle import bp_oracle
    import concurrent.futures as cf
    probe = time.perf_counter()
    bp_oracle.run_rigid(warmed[0], steps)
    per_world = max(time.perf_counter() - probe, 1e-4)
    count = int(min(len(warmed), max(threads, budget_s * threads / per_world)))
    sample = warmed[:count]
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as pool:
        list(pool.map(lambda sc: bp_oracle.run_rigid(sc, steps), sample))
    elapsed = time.perf_counter() - t0
    bodies = sum(len(sc.bodies) for sc in sample)
    return {"value": bodies * steps / elapsed, "worlds": count, "bodies": bodies, "threads": threads, "seconds": elapsed}


def cpu_batch(scenes_list, budget_s, threads, warmup, steps):
    """The oracle (FP32) on a random sample of the batch's worlds with the GPU line's schedule:
    first every sampled world's `warmup` untimed steps (worlds are added to the sample until the
    warm-up share of the wall budget is used), then the timed `steps` steps of all of them, the
    worlds spread over `threads` host threads (the oracle runs outside the interpreter lock).
    value = bubble-substeps of successful timed steps / wall time of the timed part."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import bp_oracle
    bp_oracle.load()
    bp_oracle.set_threads(1)
    order = np.random.Generator(np.random.PCG64(7)).permutation(len(scenes_list))
    deadline = time.perf_counter() + budget_s * warmup / max(warmup + steps, 1)

    def warm(w):
        if time.perf_counter() >= deadline:
            return None
        s = scenes_list[w].copy()
        if warmup:
            s.bubbles, s.bodies, _, _ = bp_oracle.run(s, warmup, assert_mode=1)
        return s

    with ThreadPoolExecutor(max_workers=threads) as pool:
        sample = [s for s in pool.map(warm, order) if s is not None]

    def timed(s):
        _, _, frames, _ = bp_oracle.run(s, steps, assert_mode=1)
        ok = frames["status"] == 0
        return len(s.bubbles) * int(frames["substeps"][ok].sum())

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as pool:
        work = sum(pool.map(timed, sample))
    wall = time.perf_counter() - t0
    return {"value": work / wall if wall > 0 else float("nan"), "worlds": len(sample), "seconds": wall,
            "bubbles": int(sum(len(s.bubbles) for s in sample)), "threads": threads, "warmup": warmup, "steps": steps}


def cpu_single(scene, steps, threads, warmup=3):
    """One world in the oracle: `warmup` untimed steps, then `steps` timed steps with `threads`
    host threads for its per-bubble phases and candidate rows (1 = sequential)."""
    from oracle import bp_oracle
    b, bod, _, _ = bp_oracle.run(scene, warmup, assert_mode=1)
    start = scene.copy()
    start.bubbles, start.bodies = b, bod
    bp_oracle.set_threads(threads)
    try:
        t = time.perf_counter()
        _, _, frames, _ = bp_oracle.run(start, steps, assert_mode=1)
        dt = time.perf_counter() - t
    finally:
        bp_oracle.set_threads(1)
    ok = frames["status"] == 0
    work = len(scene.bubbles) * int(frames["substeps"][ok].sum())
    return {"value": work / dt, "steps": int(len(frames)), "threads": threads, "seconds": dt,
            "ms_per_step": dt / max(len(frames), 1) * 1e3}


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
    ap.add_argument("--plain-steps", type=int, default=120, help="timed frames of the 256-bubble plain world")
    ap.add_argument("--plain2-steps", type=int, default=300, help="timed frames of the 1,024-bubble plain world")
    ap.add_argument("--latency-steps", type=int, default=10)
    ap.add_argument("--large-steps", type=int, default=5)
    ap.add_argument("--no-ab", action="store_true", help="skip the unpooled, graph and autotuned lines")
    ap.add_argument("--tune-repeats", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-evict-seconds", type=float, default=10.0)
    ap.add_argument("--serve-evict-budget", type=float, default=0.9, help="memory budget of the eviction run, "
                    "as a fraction of the resident footprint")
    ap.add_argument("--serve-rate", type=float, default=20000.0)
    ap.add_argument("--serve-worlds", type=int, default=8192)
    ap.add_argument("--serve-c4-seconds", type=float, default=4.0)
    ap.add_argument("--serve-c4-rates", type=float, nargs="*", default=[20000.0, 40000.0, 80000.0, 160000.0])
    ap.add_argument("--serve-streams", type=int, default=32)
    ap.add_argument("--serve-inflight", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=min(os.cpu_count() or 1, HOST_THREADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=0, help="only run the copy calibration kernel of this size")
    ap.add_argument("--rigid-worlds", type=int, default=4096, help="rigid worlds without bubbles (0 skips the line)")
    ap.add_argument("--rigid-bodies", type=int, default=48, help="dynamic bodies per rigid world")
    ap.add_argument("--rigid-steps", type=int, default=20)
    ap.add_argument("--rigid-warmup", type=int, default=120, help="untimed rigid steps before the timed ones")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the worlds are independent, so a multi-GPU run would shard them by "
              f"world with no exchange; only one MI355X is measured (--gpus {args.gpus} is not supported)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_bytes:
        copy_calibration(args.copy_bytes)
        return

    from bubblephysics_amd import scenes
    t_build = time.perf_counter()
    scenes_list = scenes.config4(worlds=args.worlds)
    log(f"built {len(scenes_list)} worlds in {time.perf_counter() - t_build:.1f}s")
    res = run_batched(args, scenes_list, mode="graph", label="batched (graph replay)")
    ab = None
    if not args.no_ab:
        line = lambda r: {"value": r["substep_bubbles"] / r["elapsed"], "ms_per_step": r["elapsed"] / args.steps * 1e3}
        ab = {"direct_launches": line(run_batched(args, scenes_list, mode="direct", label="direct launches")),
              "unpooled": line(run_batched(args, scenes_list, mode="unpooled", label="unpooled")),
              "autotuned": run_autotuned(args, scenes_list)}
    c1, c2 = scenes.config1(), scenes.config2()
    plain = run_plain(c1, args.plain_steps) if args.plain_steps > 0 else None
    plain2 = run_plain(c2, args.plain2_steps) if args.plain2_steps > 0 else None
    plain_value = plain["value"] if plain else float("nan")
    log(f"plain: {plain_value:.4g} (config 1), {plain2['value'] if plain2 else float('nan'):.4g} (config 2)")
    large = run_large(args.large_steps) if args.large_steps > 0 else None
    if large:
        log(f"large done: {large['value']:.3g} bubble-substeps/s")
    rigid_line = None
    if args.rigid_worlds > 0:
        t_build = time.perf_counter()
        rigid_list = rigid_scenes(args.rigid_worlds, args.rigid_bodies)
        log(f"built {len(rigid_list)} rigid worlds in {time.perf_counter() - t_build:.1f}s")
        rigid_line, rigid_warmed = run_rigid(args, rigid_list)
        log(f"rigid: {rigid_line['value']:.3g} body-steps/s, {rigid_line['gpu_ms']:.2f} ms per {args.rigid_steps}-step launch")
        if not args.no_cpu:
            rigid_line["cpu"] = cpu_rigid(rigid_warmed, args.cpu_threads, args.rigid_steps, args.cpu_seconds / 3)
            log(f"rigid cpu: {rigid_line['cpu']['value']:.3g} body-steps/s")
    served_c4 = run_served_sweep(args, scenes_list, plain_value) if args.serve_c4_seconds > 0 and args.serve_c4_rates else None
    served, evict = run_served(args) if args.serve_seconds > 0 else (None, None)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_batch(scenes_list, args.cpu_seconds, args.cpu_threads, args.warmup, args.steps)
        log(f"cpu batch: {cpu['value']:.3g} bubble-substeps/s on {cpu['threads']} threads")
        cpu["single"] = {
            "config1_1_thread": cpu_single(c1, args.plain_steps, 1),
            "config1_per_phase": cpu_single(c1, args.plain_steps, args.cpu_threads),
            "config2_1_thread": cpu_single(c2, 60, 1),
            "config2_per_phase": cpu_single(c2, 60, args.cpu_threads),
        }
        cpu["drift"] = fp64_drift()
        log("cpu single-world and drift done")

    value = res["substep_bubbles"] / res["elapsed"]
    # The dominant kernel: the class that moves the most algorithmic bytes in the timed launch.
    k = int(np.argmax(res["class_bytes"]))
    dur_s = float(res["class_ms"][k]) / 1e3          # one launch of the class covers all the timed steps
    achieved = res["class_bytes"][k] / dur_s / 1e9 if dur_s > 0 else 0.0
    kname = kernel_name(k, res["class_threads"], res["class_capacity"])
    traffic = recorded_traffic(kname, args.steps, res["worlds"])
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
            "worlds": res["worlds"], "bubbles": res["bubbles"], "parallelism": "one workgroup per world, pooled per class",
        },
        "p50_step_ms": percentile(res["latencies"], 50),     # one step of every world, host clock
        "p99_step_ms": percentile(res["latencies"], 99),
        "failed_worlds": res["failed_worlds"],
        "candidate_overflow_max": res["overflow"],
        "class_ms": {kernel_name(j, res["class_threads"], res["class_capacity"]): float(res["class_ms"][j])
                     for j in range(len(res["class_ms"])) if res["class_ms"][j] > 0},
        "ab": ab,
        "plain": None if plain is None else {"workload": "plain: config 1, 1 world of 256 bubbles, walls, gravity, "
                                             "one launch per frame", **plain},
        "plain_config2": None if plain2 is None else {"workload": "plain: config 2, 1 world of 1,024 bubbles + 64 kinematic "
                                                      "bodies, one launch per frame", **plain2},
        "batched_over_plain": value / plain_value if plain_value > 0 else None,
        "batched_over_plain_config2": value / plain2["value"] if plain2 and plain2["value"] > 0 else None,
        "large": None if large is None else {
            "workload": f"config 3: 1 world, {large['bubbles']} bubbles + {large['bodies']} bodies (kinematic and walls), "
                        "cooperative grid", "value": large["value"], "ms_per_step": large["ms_per_step"],
            "gpu_ms_per_step": large["gpu_ms_per_step"], "statuses": large["statuses"]},
        "rigid": None if rigid_line is None else {
            "workload": f"rigid worlds without bubbles: {rigid_line['worlds']} seeded crowds of {args.rigid_bodies} dynamic "
                        f"bodies (boxes, circles, capsules, convex polygons) + 6 statics and kinematics, 5 joints, gravity; "
                        f"{args.rigid_warmup} untimed steps (the crowds fall and pile up), then {args.rigid_steps} "
                        f"timed steps in one launch, one workgroup per world",
            "metric": "rigid body-steps/s", "value": rigid_line["value"], "unit": "body-steps/s",
            "ms_per_step": rigid_line["elapsed"] / args.rigid_steps * 1e3,
            "kernel_ms_per_launch": rigid_line["gpu_ms"], "bodies": rigid_line["bodies"], "ok_worlds": rigid_line["ok_worlds"],
            "candidate_pairs_per_step": rigid_line["pairs_per_step"], "contacts_per_world": rigid_line["contacts"],
            "cpu_baseline": None if "cpu" not in rigid_line else {
                "value": rigid_line["cpu"]["value"], "unit": "body-steps/s", "cores": rigid_line["cpu"]["threads"],
                "kind": "port", "sample": f"oracle FP32 rigid solver on the first {rigid_line['cpu']['worlds']} worlds from "
                f"their warmed-up states, {args.rigid_steps} timed steps each, over {rigid_line['cpu']['threads']} threads"}},
        "served_batch_worlds": None if served_c4 is None else {
            "workload": f"serving the {res['worlds']} batched worlds above: Poisson step requests at rising offered loads, "
                        f"{args.serve_c4_seconds:.0f}s each, 10% with transactions; over_plain against config 1's plain step",
            **served_c4},
        "served": None if served is None else {
            "workload": f"serving: {served['worlds']} resident worlds, bubbles log-uniform in [32, 16384] "
                        f"({served['bubbles']} total), Poisson {args.serve_rate:.0f} step requests/s for "
                        f"{args.serve_seconds:.0f}s, 10% with transactions, budget 200 GB",
            **{kk: v for kk, v in served.items() if kk not in ("worlds", "bubbles")},
            "under_eviction": evict},
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
            "kernel": kname,
            "kernel_ms_per_launch": float(res["class_ms"][k]),
        },
        "cpu_baseline": None if cpu is None else {
            "value": cpu["value"], "unit": "bubble-substeps/s", "cores": cpu["threads"], "kind": "port",
            "sample": f"oracle FP32 on {cpu['worlds']} randomly chosen worlds ({cpu['bubbles']} bubbles) of the same "
                      f"batch, each {cpu['warmup']} untimed + {cpu['steps']} timed steps (the GPU line's schedule), "
                      f"worlds spread over {cpu['threads']} host threads ({cpu['seconds']:.1f}s wall for the timed part)",
            "host": {"nproc": os.cpu_count(), "model": cpu_model(), "threads_used": cpu["threads"]},
            "single_world": cpu["single"],
            "fp64_drift": cpu["drift"],
        },
    }
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
