"""Serving: step and transaction requests for many resident worlds, by continuous batching.

Requests arrive as a seeded Poisson process; each asks for one step of one world, some carry a
transaction applied before that step (impulse, add, remove, merge or split of bubbles). The loop
keeps one launch in flight: every world with a waiting request contributes its oldest request to
the next launch, and requests that arrive meanwhile wait for the launch after it. Worlds are
admitted against a device-memory budget; the least recently used idle worlds are evicted to host
memory when an admission needs room, and come back on their next request.

This has no counterpart in the reference, which steps one world per call. The transaction
semantics follow the reference's staging in simplified form (TransactionStager.swift:312-324 for
impulses, StructuralOperations.swift:10-151 for merge and split: area, mass and momentum kept).
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi, scenes, solver

KINDS = ("impulse", "add", "remove", "merge", "split")
BIG_WORLD = 4096          # above this a world runs in the one-workgroup global kernel while serving


@dataclass
class Request:
    arrival: float         # seconds after the start of the run
    world: int
    transaction: str | None
    seed: int
    done: float = math.nan
    status: int = -1
    substeps: int = 0


def poisson_requests(worlds: int, rate: float, duration: float, seed: int, transaction_fraction: float = 0.1):
    """Arrivals of a Poisson process of `rate` requests/s over `duration` s, worlds uniform."""
    rng = np.random.Generator(np.random.PCG64(seed))
    count = int(rng.poisson(rate * duration))
    arrivals = np.sort(rng.uniform(0.0, duration, count))
    world = rng.integers(0, worlds, count)
    with_tx = rng.uniform(0.0, 1.0, count) < transaction_fraction
    kind = rng.integers(0, len(KINDS), count)
    seeds = rng.integers(0, 2**31 - 1, count)
    return [Request(float(arrivals[k]), int(world[k]), KINDS[int(kind[k])] if with_tx[k] else None, int(seeds[k]))
            for k in range(count)]


def apply_transaction(bubbles: np.ndarray, kind: str, seed: int) -> np.ndarray:
    """The world's bubbles after one transaction (float32 arithmetic, deterministic in `seed`)."""
    b = bubbles.copy()
    n = len(b)
    if n == 0:
        return b
    rng = np.random.Generator(np.random.PCG64(seed))
    i = int(rng.integers(0, n))
    f32 = np.float32
    if kind == "impulse":                                   # v += J / m
        if b["inverse_mass"][i] > 0:
            j = rng.normal(0.0, 50.0, 2).astype(f32) * b["mass"][i]
            b["velocity"][i] = (b["velocity"][i] + j * b["inverse_mass"][i]).astype(f32)
        return b
    if kind == "remove" and n > 1:
        return np.delete(b, i)
    if kind == "add":
        new = b[i:i + 1].copy()
        new["id"] = int(b["id"].max()) + 1
        angle = f32(rng.uniform(0.0, 2.0 * math.pi))
        offset = f32(2.0) * b["radius"][i]
        new["center"][0] = b["center"][i] + offset * np.array([math.cos(angle), math.sin(angle)], f32)
        new["previous_center"] = new["center"]
        return np.concatenate([b, new])
    if kind == "merge" and n > 1:                            # two neighbours by index become one
        k = i + 1 if i + 1 < n else i - 1
        a, c = b[min(i, k)], b[max(i, k)]
        m = f32(a["mass"] + c["mass"])
        merged = b[min(i, k):min(i, k) + 1].copy()
        merged["radius"] = f32(math.sqrt(float(a["radius"]) ** 2 + float(c["radius"]) ** 2))
        merged["mass"] = m
        merged["inverse_mass"] = f32(1.0) / m if m > 0 else f32(0.0)
        merged["center"] = ((a["center"] * a["mass"] + c["center"] * c["mass"]) / m).astype(f32)
        merged["previous_center"] = merged["center"]
        merged["velocity"] = ((a["velocity"] * a["mass"] + c["velocity"] * c["mass"]) / m).astype(f32)
        keep = np.delete(b, [min(i, k), max(i, k)])
        return np.sort(np.concatenate([keep, merged]), order="id")
    if kind == "split":                                      # one bubble becomes two halves
        a = b[i]
        half = b[i:i + 1].copy()
        half["radius"] = f32(a["radius"] / math.sqrt(2.0))
        half["mass"] = f32(a["mass"] * f32(0.5))
        half["inverse_mass"] = f32(1.0) / half["mass"] if half["mass"][0] > 0 else f32(0.0)
        angle = f32(rng.uniform(0.0, 2.0 * math.pi))
        d = (np.array([math.cos(angle), math.sin(angle)], f32) * half["radius"][0]).astype(f32)
        one, two = half.copy(), half.copy()
        one["center"] = (a["center"] + d).astype(f32)
        two["center"] = (a["center"] - d).astype(f32)
        two["id"] = int(b["id"].max()) + 1
        for x in (one, two):
            x["previous_center"] = x["center"]
        keep = np.delete(b, i)
        return np.sort(np.concatenate([keep, one, two]), order="id")
    return b


def serving_worlds(count: int, seed: int, size=(32, 16384)):
    """Worlds of the serving configuration: bubble counts log-uniform in `size`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.exp(rng.uniform(math.log(size[0]), math.log(size[1]), count)).astype(np.int64)
    return [scenes.batch_world(rng, int(n), name="serve") for n in sizes]


@dataclass
class ServeStats:
    served: int = 0
    bubble_substeps: int = 0
    elapsed: float = 0.0
    launches: int = 0
    evictions: int = 0
    admissions: int = 0
    failed: int = 0
    transaction_seconds: float = 0.0   # host time spent applying transactions
    launch_sizes: list = field(default_factory=list)
    latencies_ms: list = field(default_factory=list)

    def summary(self) -> dict:
        lat = np.asarray(self.latencies_ms, np.float64)
        return {
            "served": self.served, "launches": self.launches, "elapsed_s": self.elapsed,
            "requests_per_s": self.served / self.elapsed if self.elapsed else 0.0,
            "bubble_substeps_per_s": self.bubble_substeps / self.elapsed if self.elapsed else 0.0,
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else math.nan,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else math.nan,
            "evictions": self.evictions, "admissions": self.admissions, "failed_steps": self.failed,
            "mean_worlds_per_launch": float(np.mean(self.launch_sizes)) if self.launch_sizes else 0.0,
            "transaction_seconds": self.transaction_seconds,
        }


class Server:
    """Continuous batching over a bp_batch of worlds with a device-memory budget (bytes)."""

    def __init__(self, worlds, *, memory_budget: int):
        self.batch = solver.Batch(worlds)
        self.counts = [len(w.bubbles) for w in worlds]
        for w, n in enumerate(self.counts):
            if n > BIG_WORLD:
                self.batch.set_world_class(w, 10)   # concurrent with the other classes
        self.budget = memory_budget
        self.resident = collections.OrderedDict()   # world -> footprint, least recently used first
        self.stats = ServeStats()
        self.used = 0
        # Worlds start resident while they fit the budget, in index order; the rest wait in host memory.
        for w in range(len(worlds)):
            need = self.batch.footprint(w)
            if self.used + need <= self.budget:
                self.resident[w] = need
                self.used += need
            else:
                self.batch.set_resident(w, False)

    def _admit(self, world: int, busy: set):
        if world in self.resident:
            self.resident.move_to_end(world)
            return
        need = self.batch.footprint(world)
        while self.used + need > self.budget:
            victim = next((w for w in self.resident if w not in busy), None)
            if victim is None:
                raise RuntimeError("memory budget too small for the worlds of one launch")
            self.used -= self.resident.pop(victim)
            self.batch.set_resident(victim, False)
            self.stats.evictions += 1
        self.batch.set_resident(world, True)
        self.resident[world] = need
        self.used += need
        self.stats.admissions += 1

    def run(self, requests, *, on_done=None) -> ServeStats:
        """Serves every request (arrival times relative to the call) and returns the statistics."""
        queues = collections.defaultdict(collections.deque)
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        in_flight = None
        t0 = time.perf_counter()
        while pending or in_flight is not None or any(queues.values()):
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                queues[r.world].append(r)
            if in_flight is None:
                ready = [w for w, q in queues.items() if q]
                if not ready:
                    if pending:
                        time.sleep(min(max(pending[0].arrival - now, 0.0), 0.001))
                    continue
                ready.sort(key=lambda w: queues[w][0].arrival)
                batch_list = [(w, queues[w].popleft()) for w in ready]
                busy = {w for w, _ in batch_list}
                for w, r in batch_list:
                    self._admit(w, busy)
                t_tx = time.perf_counter()
                for w, r in batch_list:
                    if r.transaction:
                        nb = apply_transaction(self.batch.read_bubbles(w), r.transaction, r.seed)
                        self.batch.write_bubbles(w, nb)
                        self.counts[w] = len(nb)
                self.stats.transaction_seconds += time.perf_counter() - t_tx
                self.batch.submit_worlds([w for w, _ in batch_list], 1)
                self.stats.launch_sizes.append(len(batch_list))
                in_flight = batch_list
                self.stats.launches += 1
                continue
            infos = self.batch.poll()
            if infos is None:
                time.sleep(0.0001)
                continue
            done = time.perf_counter() - t0
            for j, (w, r) in enumerate(in_flight):
                info = infos[0, j]
                r.done, r.status, r.substeps = done, int(info["status"]), int(info["substep_count"])
                self.stats.latencies_ms.append((done - r.arrival) * 1e3)
                if r.status == abi.BP_OK:                      # work of failed steps is not counted
                    self.stats.bubble_substeps += self.counts[w] * r.substeps
                self.stats.served += 1
                self.stats.failed += r.status != abi.BP_OK
                if on_done:
                    on_done(r)
            in_flight = None
        self.stats.elapsed = time.perf_counter() - t0
        return self.stats
