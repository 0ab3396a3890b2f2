"""Autotuner of a batch's kernel configuration per world size (bp_batch_autotune).

For each bin of world sizes the library times one step of the bin's worlds in every kernel
configuration that holds them (workgroup width x bubbles per lane of the LDS-resident kernel, and
the global-state kernel: the LDS-to-HBM cutover), then the best one with half and twice the
automatic grid bucket count, from the batch's published state, which it leaves unchanged. The
table can be cached per GPU architecture and applied to later batches without trials. The
choice never changes results, only time.
"""
from __future__ import annotations

import ctypes
import json
import os
from pathlib import Path

import numpy as np

from . import abi


def tune(batch, repeats: int = 3) -> np.ndarray:
    """Runs the trials on `batch` and adopts the choices; returns the table (TUNE_DTYPE records)."""
    lib = abi.load_library()
    out = np.zeros(32, abi.TUNE_DTYPE)
    n = ctypes.c_uint32(0)
    abi.check(lib.bp_batch_autotune(batch._handle, repeats, abi.ptr(out), len(out), ctypes.byref(n)), "bp_batch_autotune")
    return out[: n.value].copy()


def apply(batch, table: np.ndarray):
    """Adopts a table (from tune or from a cache) without trials."""
    t = np.ascontiguousarray(table, abi.TUNE_DTYPE)
    abi.check(abi.load_library().bp_batch_apply_tuning(batch._handle, abi.ptr(t), len(t)), "bp_batch_apply_tuning")


def to_json(table: np.ndarray, arch: str) -> dict:
    return {"arch": arch, "bins": [{"min_bubbles": int(r["min_bubbles"]), "max_bubbles": int(r["max_bubbles"]),
                                    "best_class": int(r["best_class"]), "best_bucket_shift": int(r["best_bucket_shift"]),
                                    "best_rate": float(r["best_rate"]), "worlds": int(r["worlds"]),
                                    "class_rate": [float(x) for x in r["class_rate"]],
                                    "bucket_rate": [float(x) for x in r["bucket_rate"]]} for r in table]}


def from_json(doc: dict) -> np.ndarray:
    t = np.zeros(len(doc["bins"]), abi.TUNE_DTYPE)
    for k, b in enumerate(doc["bins"]):
        for name in ("min_bubbles", "max_bubbles", "best_class", "best_bucket_shift", "best_rate", "worlds"):
            t[k][name] = b[name]
        t[k]["class_rate"] = b["class_rate"]
        t[k]["bucket_rate"] = b["bucket_rate"]
    return t


def cache_path(arch: str) -> Path:
    root = Path(os.environ.get("BUBBLEPHYSICS_TUNING_DIR", Path.home() / ".cache" / "bubblephysics_amd"))
    return root / f"tuning-{arch}.json"


def tune_cached(batch, arch: str = "gfx950", repeats: int = 3, path: os.PathLike | None = None) -> np.ndarray:
    """Applies the cached table of `arch` when one exists, else tunes and writes the cache."""
    p = Path(path) if path is not None else cache_path(arch)
    if p.exists():
        table = from_json(json.loads(p.read_text()))
        apply(batch, table)
        return table
    table = tune(batch, repeats)
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(json.dumps(to_json(table, arch), indent=1))
    return table
