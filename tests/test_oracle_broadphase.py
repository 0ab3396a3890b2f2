"""Pins the oracle's broadphase: candidate rows equal the k smallest keys of a brute-force cell
neighbourhood (the fixture of Tests/BubblePhysicsNativeMetalTests/NativeMetalChainKernelTests
.swift:13-20, 507-600, with its portable LCG), for the sorted-grid and the hashed-grid emission,
including truncated rows and their overflow counters."""
import numpy as np
import pytest

from bubblephysics_amd import abi, scenes

MASK64 = (1 << 64) - 1


class Lcg:
    """The portable generator of NativeMetalChainKernelTests.swift:13-20."""

    def __init__(self, state):
        self.state = state & MASK64

    def next(self):
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & MASK64
        return (self.state >> 33) & 0xFFFFFFFF

    def unit(self):
        return np.float32(self.next() % 100000) / np.float32(100000)


def fixture(count, limit, spread):
    gen = Lcg(count * 131 + limit)
    spread = np.float32(spread)
    recs = []
    for index in range(count):
        radius = np.float32(0.2) + gen.unit() * np.float32(0.7)
        x = gen.unit() * spread - spread / np.float32(2)
        y = gen.unit() * spread - spread / np.float32(2)
        ident = 100 + (index * 7919) % 100003
        filt = (1, 0) if index % 9 == 4 else (1, 0xFFFFFFFF)
        recs.append((ident, radius, x, y, filt))
    recs.sort(key=lambda r: r[0])
    b = np.zeros(count, abi.BUBBLE_DTYPE)
    for k, (ident, radius, x, y, filt) in enumerate(recs):
        b[k]["id"] = ident
        b[k]["radius"] = radius
        b[k]["center"] = (x, y)
        b[k]["previous_center"] = (x, y)
        b[k]["mass"] = 1
        b[k]["inverse_mass"] = 1
        b[k]["flags"] = abi.BP_BUBBLE_ENABLED
        b[k]["category"], b[k]["mask"] = filt
    segs = np.zeros(5, abi.SEGMENT_DTYPE)
    for j in range(5):
        segs[j]["id"] = j + 1
        segs[j]["a"] = (gen.unit() * spread - spread / np.float32(2), gen.unit() * spread - spread / np.float32(2))
        segs[j]["b"] = (gen.unit() * spread - spread / np.float32(2), gen.unit() * spread - spread / np.float32(2))
        segs[j]["category"], segs[j]["mask"] = 1, 0xFFFFFFFF
    return b, segs


def brute_force(b, segs, cell, slack, limit):
    """Every pair in adjacent cells that passes the filter and the Chebyshev reach, then segments."""
    def cell_of(p):
        return (int(np.floor(np.float32(p[0]) / np.float32(cell))), int(np.floor(np.float32(p[1]) / np.float32(cell))))

    def accepts(ac, am, bc, bm):
        return (ac & bm) != 0 and (bc & am) != 0

    rows, dropped, widest = [], 0, 0
    for i in range(len(b)):
        a = b[i]
        keys = []
        ca = cell_of(a["center"])
        for j in range(i + 1, len(b)):
            o = b[j]
            cb = cell_of(o["center"])
            if abs(ca[0] - cb[0]) > 1 or abs(ca[1] - cb[1]) > 1:
                continue
            if not accepts(a["category"], a["mask"], o["category"], o["mask"]):
                continue
            reach = np.float32(np.float32(a["radius"] + o["radius"]) + np.float32(slack))
            if abs(np.float32(a["center"][0] - o["center"][0])) <= reach and abs(np.float32(a["center"][1] - o["center"][1])) <= reach:
                keys.append((0, int(a["id"]), int(o["id"]), j))
        for j, s in enumerate(segs):
            if not accepts(a["category"], a["mask"], s["category"], s["mask"]):
                continue
            reach = np.float32(a["radius"] + np.float32(slack))
            lo = np.minimum(s["a"], s["b"]) - reach
            hi = np.maximum(s["a"], s["b"]) + reach
            if lo[0] <= a["center"][0] <= hi[0] and lo[1] <= a["center"][1] <= hi[1]:
                keys.append((1, int(a["id"]), int(s["id"]), j))
        keys.sort()
        dropped += max(0, len(keys) - limit)
        widest = max(widest, len(keys))
        rows.append(keys[:limit])
    return rows, dropped, widest


@pytest.mark.parametrize("count,limit,spread", [(1, 8, 4), (60, 8, 12), (200, 64, 30), (200, 6, 14), (301, 32, 20)])
@pytest.mark.parametrize("hashed,buckets", [(False, 2048), (True, 2048), (True, 1), (True, 7)])
def test_candidate_rows_match_the_cell_neighbourhood_reference(oracle, count, limit, spread, hashed, buckets):
    b, segs = fixture(count, limit, spread)
    cell, slack = np.float32(2.0), np.float32(0.05)
    rows, counts, overflow = oracle.candidate_rows(b, np.zeros(count, np.float32), segs, cell, slack, limit,
                                                   hashed=hashed, buckets=buckets)
    expected, dropped, widest = brute_force(b, segs, cell, slack, limit)
    non_empty = 0
    for i, keys in enumerate(expected):
        assert int(counts[i].sum()) == len(keys), f"bubble {i}"
        assert int(counts[i, 0]) == sum(1 for k in keys if k[0] == 0)
        for slot, key in enumerate(keys):
            assert (int(rows[i, slot, 0]), int(rows[i, slot, 1])) == (key[0], key[3]), f"bubble {i} slot {slot}"
        non_empty += bool(keys)
    assert int(overflow[0]) == dropped
    assert int(overflow[1]) == widest
    if count >= 60:
        assert non_empty > 5
    if limit == 6:
        assert dropped > 0


def test_sorted_and_hashed_grids_step_identical_worlds(oracle):
    scene = scenes.config2(steps=3)
    b1, _, f1, x1 = oracle.run(scene, 3, grid="sorted", adjacency=True)
    b2, _, f2, x2 = oracle.run(scene, 3, grid="hashed", adjacency=True)
    assert np.array_equal(b1["center"].view(np.uint32), b2["center"].view(np.uint32))
    assert np.array_equal(x1["entries"], x2["entries"])
    assert np.array_equal(f1["substeps"], f2["substeps"])


def test_adjacency_rows_are_sorted_and_symmetric(oracle):
    scene = scenes.config1(steps=2)
    _, _, _, ex = oracle.run(scene, 2, adjacency=True)
    off, ent = ex["row_offsets"], ex["entries"]
    pairs = set()
    for i in range(len(scene.bubbles)):
        row = ent[off[i]:off[i + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0), "rows ordered by (kind, neighbour) without duplicates"
        for e in row:
            if e >> 30 == 0:
                pairs.add((i, int(e & ((1 << 30) - 1))))
    assert all((j, i) in pairs for i, j in pairs), "a pair sits in the rows of both bubbles"
