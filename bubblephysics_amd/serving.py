"""Serving: step and transaction requests for many resident worlds, by continuous batching.

The request loop is native (bp_server in the HIP library): a dispatcher thread takes requests
as they arrive, launches per kernel class the waiting worlds that are not already on the
device, keeps several launches in flight on several streams and posts each result when its
launch completes. Transactions run on the device before the step they come with. Worlds are
admitted against a device-memory budget; the least recently used idle worlds are evicted to
host memory when an admission needs room.

This module builds request traces (a seeded Poisson process), restates how a served transaction
changes a world (``apply_transaction``: the same rules, for replaying a world's requests alone) and
wraps the server. The reference has no serving form: its game loop steps one world per call
(BubbleWorld.submit/advance, Sources/BubblePhysicsPublic/BubbleWorld.swift:252-283). The
transactions follow its staging: the impulse with its role and finiteness checks
(TransactionStager.swift:312-324, in FP64), add with the placement rules (:406-421,
PlacementGeometry.swift:42-94), merge and split with StructuralOperations.swift:10-151.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import abi, scenes, solver

KINDS = ("impulse", "add", "remove", "merge", "split")
TX_CODE = {None: abi.BP_TX_NONE, "impulse": abi.BP_TX_IMPULSE, "add": abi.BP_TX_ADD, "remove": abi.BP_TX_REMOVE,
           "merge": abi.BP_TX_MERGE, "split": abi.BP_TX_SPLIT}


@dataclass
class Request:
    arrival: float                   # seconds after the start of the trace
    world: int
    transaction: str | None = None
    index: int = 0                   # bubble of the transaction: index % bubble count
    impulse: tuple = (0.0, 0.0)      # impulse J of an impulse transaction (velocity change J / m)
    direction: tuple = (1.0, 0.0)    # unit vector of add and split
    # Filled by the server.
    done: float = math.nan
    status: int = -1
    substeps: int = 0
    bubbles: int = 0
    tx_result: int = 0               # BP_REJECT_* of its transaction (0 = applied or none)


def transaction_params(rng, kind: str | None) -> dict:
    """Random parameters of one transaction, rounded to FP32 as the device receives them."""
    angle = float(rng.uniform(0.0, 2.0 * math.pi))
    return {"transaction": kind, "index": int(rng.integers(0, 2**31 - 1)),
            "impulse": tuple(float(x) for x in rng.normal(0.0, 50.0, 2).astype(np.float32)),
            "direction": (float(np.float32(math.cos(angle))), float(np.float32(math.sin(angle))))}


def poisson_requests(worlds: int, rate: float, duration: float, seed: int, transaction_fraction: float = 0.1):
    """Arrivals of a Poisson process of `rate` requests/s over `duration` s, worlds uniform; a
    `transaction_fraction` of them carry a transaction, the five kinds equally likely."""
    rng = np.random.Generator(np.random.PCG64(seed))
    count = int(rng.poisson(rate * duration))
    arrivals = np.sort(rng.uniform(0.0, duration, count))
    world = rng.integers(0, worlds, count)
    with_tx = rng.uniform(0.0, 1.0, count) < transaction_fraction
    kind = rng.integers(0, len(KINDS), count)
    out = []
    for k in range(count):
        p = transaction_params(rng, KINDS[int(kind[k])] if with_tx[k] else None)
        out.append(Request(float(arrivals[k]), int(world[k]), **p))
    return out


def apply_transaction(bubbles: np.ndarray, req: Request, *, bodies=None, vertices=None, region=None, bands=None):
    """(bubbles, reason) after one served transaction under the reference's rules, as the server
    applies them: the impulse and the removal as the device kernel states them, add, merge and
    split through the library's host staging (bp_stage_add, bp_stage_merge, bp_stage_split) with
    the world's bodies at their current poses. `reason` is BP_REJECT_* (0 = applied); a refused
    transaction returns the bubbles unchanged."""
    b = bubbles.copy()
    n = len(b)
    if n == 0 or req.transaction is None:
        return b, abi.BP_REJECT_NONE
    f32 = np.float32
    i = int(req.index) % n
    kind = req.transaction
    if kind == "impulse":                       # TransactionStager.swift:312-324
        J = np.array(req.impulse, f32)
        if not np.all(np.isfinite(J)):
            return b, abi.BP_REJECT_INVALID_VALUE
        banded = bands is not None and int(b["id"][i]) in set(int(x) for x in bands["bubble_id"])
        if not (b["flags"][i] & abi.BP_BUBBLE_FORCE_DRIVEN) or (b["flags"][i] & abi.BP_BUBBLE_STATIC) or banded:
            return b, abi.BP_REJECT_INCOMPATIBLE_MOTION_ROLE
        m = np.float64(b["mass"][i])
        with np.errstate(over="ignore", invalid="ignore"):
            v = (b["velocity"][i].astype(np.float64) + J.astype(np.float64) / m).astype(f32)
        if not np.all(np.isfinite(v)):
            return b, abi.BP_REJECT_INVALID_VALUE
        b["velocity"][i] = v
        return b, abi.BP_REJECT_NONE
    if kind == "remove":
        return (np.delete(b, i) if n > 1 else b), abi.BP_REJECT_NONE
    kw = dict(bodies=bodies, vertices=vertices, region=tuple(region) if region is not None else (-1e30, -1e30, 1e30, 1e30))
    new_id = int(b["id"][n - 1]) + 1
    try:
        if kind == "add":
            rec = b[i].copy()
            rec["id"] = new_id
            offset = f32(2.0) * b["radius"][i]
            direction = np.array(req.direction, f32)
            rec["center"] = b["center"][i] + offset * direction
            rec["previous_center"] = rec["center"]
            rec["velocity"] = 0.0
            rec["rotation"] = 0.0
            rec["angular_velocity"] = 0.0
            return solver.stage_add(b, rec, **kw), abi.BP_REJECT_NONE
        if kind == "merge":
            if n < 2:
                return b, abi.BP_REJECT_INVALID_STRUCTURAL_OPERATION
            k = i + 1 if i + 1 < n else i - 1
            a, c = min(i, k), max(i, k)
            return solver.stage_merge(b, [b["id"][a], b["id"][c]], b[a], **kw), abi.BP_REJECT_NONE
        if kind == "split":
            second = b[i].copy()
            second["id"] = new_id
            return solver.stage_split(b, b["id"][i], b[i], 0.5, second, 0.5, **kw), abi.BP_REJECT_NONE
    except solver.StageRejected as e:
        return b, e.reason
    raise ValueError(f"unknown transaction {kind!r}")


def to_trace(requests) -> np.ndarray:
    """bp_request records of a list of Request (tag = position)."""
    t = np.zeros(len(requests), abi.REQUEST_DTYPE)
    for k, r in enumerate(requests):
        t[k]["arrival"] = r.arrival
        t[k]["tag"] = k
        t[k]["world"] = r.world
        t[k]["transaction"] = TX_CODE[r.transaction]
        t[k]["index"] = r.index
        t[k]["impulse"] = r.impulse
        t[k]["direction"] = r.direction
    return t


def serving_worlds(count: int, seed: int, size=(32, 16384)):
    """Worlds of the serving configuration: bubble counts log-uniform in `size`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.exp(rng.uniform(math.log(size[0]), math.log(size[1]), count)).astype(np.int64)
    return [scenes.batch_world(rng, int(n), name="serve") for n in sizes]


@dataclass
class ServeStats:
    results: np.ndarray              # bp_server_result records, in trace order
    elapsed: float                   # seconds from the start of the trace to the last result
    counters: dict

    def summary(self) -> dict:
        r = self.results
        lat = (r["completed"] - r["arrival"]) * 1e3
        ok = r["status"] == abi.BP_OK
        statuses = {abi.STATUS_NAMES.get(int(s), int(s)): int(c) for s, c in zip(*np.unique(r["status"], return_counts=True))}
        work = int((r["bubbles"][ok].astype(np.int64) * r["substeps"][ok].astype(np.int64)).sum())
        return {
            "served": int(len(r)), "elapsed_s": self.elapsed,
            "requests_per_s": len(r) / self.elapsed if self.elapsed else 0.0,
            "bubble_substeps_per_s": work / self.elapsed if self.elapsed else 0.0,
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else math.nan,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else math.nan,
            "p999_ms": float(np.percentile(lat, 99.9)) if len(lat) else math.nan,
            "p99_ms_ok_steps": float(np.percentile(lat[ok], 99)) if ok.any() else math.nan,
            "statuses": statuses,
            "mean_worlds_per_launch": self.counters["launched_worlds"] / max(self.counters["launches"], 1),
            **{k: self.counters[k] for k in ("launches", "transactions", "rejected", "evictions", "admissions", "repacks",
                                              "resident_bytes")},
            "admission_s": self.counters.get("admission_microseconds", 0) * 1e-6,
        }


class Server:
    """bp_server over a batch of worlds with a device-memory budget (bytes, 0 = no limit)."""

    def __init__(self, worlds, *, memory_budget: int = 0, streams: int = 0, max_inflight_per_class: int = 0,
                 max_worlds_per_launch: int = 0, big_world: int = 0):
        self._lib = abi.load_library()
        self.batch = worlds if isinstance(worlds, solver.Batch) else solver.Batch(worlds)
        cfg = abi.ServerConfig()
        self._lib.bp_server_config_default(ctypes.byref(cfg))
        cfg.memory_budget = memory_budget
        for name, value in (("streams", streams), ("max_inflight_per_class", max_inflight_per_class),
                            ("max_worlds_per_launch", max_worlds_per_launch), ("big_world", big_world)):
            if value:
                setattr(cfg, name, value)
        self._handle = ctypes.c_void_p()
        abi.check(self._lib.bp_server_create(self.batch._handle, ctypes.byref(cfg), ctypes.byref(self._handle)),
                  "bp_server_create")

    def clock(self) -> float:
        return float(self._lib.bp_server_clock(self._handle))

    def run(self, requests) -> ServeStats:
        """Replays the requests at their arrival times (natively) and fills their results."""
        trace = to_trace(requests)
        results = np.zeros(len(trace), abi.RESULT_DTYPE)
        base = self.clock()
        abi.check(self._lib.bp_server_run_trace(self._handle, abi.ptr(trace), len(trace), abi.ptr(results)),
                  "bp_server_run_trace")
        for r, res in zip(requests, results):
            r.done, r.status, r.substeps, r.bubbles = float(res["completed"]), int(res["status"]), int(res["substeps"]), \
                int(res["bubbles"])
            r.tx_result = int(res["transaction_result"])
        elapsed = float(results["completed"].max() - base) if len(results) else 0.0
        return ServeStats(results, elapsed, self.stats())

    def submit(self, req: Request, tag: int = 0):
        rec = to_trace([req])
        rec[0]["tag"] = tag
        abi.check(self._lib.bp_server_submit(self._handle, abi.ptr(rec)), "bp_server_submit")

    def poll(self) -> np.ndarray:
        out = np.zeros(4096, abi.RESULT_DTYPE)
        n = ctypes.c_uint32(0)
        abi.check(self._lib.bp_server_poll(self._handle, abi.ptr(out), len(out), ctypes.byref(n)), "bp_server_poll")
        return out[:n.value].copy()

    def wait_idle(self):
        abi.check(self._lib.bp_server_wait_idle(self._handle), "bp_server_wait_idle")

    def read_bubbles(self, world: int) -> np.ndarray:
        n = ctypes.c_uint32(0)
        abi.check(self._lib.bp_server_read_bubbles(self._handle, world, None, 0, ctypes.byref(n)), "bp_server_read_bubbles")
        out = np.zeros(n.value, abi.BUBBLE_DTYPE)
        abi.check(self._lib.bp_server_read_bubbles(self._handle, world, abi.ptr(out), len(out), ctypes.byref(n)),
                  "bp_server_read_bubbles")
        return out

    def stats(self) -> dict:
        s = abi.ServerStats()
        abi.check(self._lib.bp_server_get_stats(self._handle, ctypes.byref(s)), "bp_server_get_stats")
        return {name: int(getattr(s, name)) for name, _ in abi.ServerStats._fields_}

    def close(self):
        """Stops the server; the batch keeps every world's newest state."""
        if self._handle:
            self._lib.bp_server_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
