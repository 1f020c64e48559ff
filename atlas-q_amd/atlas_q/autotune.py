"""Launch-shape autotuner for the grouped persistent kernels.

The grouped kernels take the number of persistent work-groups as a free
parameter, and the grouped GEMM also its tile shape (64-row tiles of 256
threads or 32-row tiles of 128 threads; a group of few-row problems wastes
half as much of each 32-row tile).  For each (kernel family, work-size
bucket) the tuner times the candidate shapes on the real work list with HIP
events (idempotent launches only), keeps the fastest, and persists the table
as JSON (``$ATLASQ_AUTOTUNE_CACHE``, default ``autotune_cache.json`` next to
this file).  Untuned keys use a heuristic: 64-row tiles, enough work-groups
for every tile up to four per CU (256 CUs) for GEMM families, and up to 16
waves per CU for element-wise families.
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path

import torch

CUS = 256
CANDIDATES = {"gemm": [256, 512, 1024, 2048], "elem": [1024, 4096, 16384]}
TILE_ROWS = (64, 32)   # candidate tile heights of the grouped GEMM
DEFAULT_PATH = Path(os.environ.get("ATLASQ_AUTOTUNE_CACHE",
                                   Path(__file__).with_name("autotune_cache.json")))


def _family(key: str) -> str:
    return "elem" if key in ("elem",) else "gemm"


def _bucket(total: int) -> int:
    return int(math.log2(max(1, total)))


class Autotuner:
    _default = None

    def __init__(self, path: Path = DEFAULT_PATH, tune: bool = True):
        self.path = Path(path)
        self.tune = tune
        self.table = {}
        if self.path.exists():
            try:
                self.table = json.loads(self.path.read_text())
            except (OSError, ValueError):
                self.table = {}
        self.trials = []

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = cls(tune=os.environ.get("ATLASQ_AUTOTUNE", "1") != "0")
        return cls._default

    def heuristic(self, key: str, total: int) -> int:
        if _family(key) == "elem":
            return max(1, min(-(-total // 256), CUS * 16))
        return max(1, min(total, 4 * CUS))

    def grid(self, key: str, total: int, trial_fn=None) -> int:
        k = f"{key}:{_bucket(total)}"
        if k in self.table:
            return int(self.table[k])
        if not (self.tune and trial_fn is not None):
            return self.heuristic(key, total)
        best, best_ms = None, float("inf")
        for g in CANDIDATES[_family(key)]:
            g = max(1, min(g, total if _family(key) == "gemm" else g))
            trial_fn(g)  # warm
            s = torch.cuda.Event(enable_timing=True)
            e = torch.cuda.Event(enable_timing=True)
            s.record()
            for _ in range(3):
                trial_fn(g)
            e.record()
            e.synchronize()
            ms = s.elapsed_time(e) / 3
            self.trials.append((k, g, ms))
            if ms < best_ms:
                best, best_ms = g, ms
        self.table[k] = best
        return best

    def shape(self, key: str, totals: dict, trial_fn=None):
        """(tile rows, grid) for a grouped GEMM whose tile count is totals[rows]
        for each candidate tile height; trial_fn(rows, grid) launches it."""
        base = totals[64]
        k = f"{key}:shape:{_bucket(base)}"
        if k in self.table:
            rows, g = self.table[k]
            return int(rows), int(g)
        if not (self.tune and trial_fn is not None):
            return 64, self.heuristic(key, base)
        best, best_ms = None, float("inf")
        for rows, total in sorted(totals.items(), reverse=True):
            if rows not in TILE_ROWS:
                continue
            for g in CANDIDATES["gemm"]:
                g = max(1, min(g, total))
                trial_fn(rows, g)  # warm
                s = torch.cuda.Event(enable_timing=True)
                e = torch.cuda.Event(enable_timing=True)
                s.record()
                for _ in range(3):
                    trial_fn(rows, g)
                e.record()
                e.synchronize()
                ms = s.elapsed_time(e) / 3
                self.trials.append((k, (rows, g), ms))
                if ms < best_ms:
                    best, best_ms = (rows, g), ms
        self.table[k] = list(best)
        return best

    def save(self, path: Path = None):
        p = Path(path) if path else self.path
        p.write_text(json.dumps(self.table, indent=1, sort_keys=True) + "\n")
