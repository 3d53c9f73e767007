"""Autotuner for the batched LU: block size and GEMM tile.

The LU block size ``nb`` sets the panel width (latency-bound work that grows
with nb) against the K depth of the MFMA trailing updates (arithmetic
intensity that grows with nb); the GEMM tile variant trades LDS reuse
against occupancy and edge waste on small systems.  The best pair depends
on the system size, the batch size, the number of right-hand sides and the
precision, so it is measured: for a shape bucket the tuner builds a
synthetic batch of sphere mode systems of about that size, times the solve
phase (HIP events on the handle's stream) for each candidate, and caches the
winner per (device, precision, n bucket, systems bucket, columns bucket) in
a JSON file.

``best`` is what ``bor_solver.solve_bor_batch`` calls when no block size is
given; ``bench.py`` calls it for the benchmark's own shape before the timed
steps.  The cache path is ``$GHOST_BOR_TUNE_CACHE``, else
``~/.cache/ghost_bor_tune.json``; ``set_cache_path`` changes it.  The device
key comes from the library's own HIP runtime (ghost_device_name).
"""

from __future__ import annotations

import json
import math
import os
import threading
from typing import Dict, Optional, Tuple

import numpy as np

from . import _lib

DEFAULT = {_lib.GHOST_FP64: (64, 4), _lib.GHOST_MIXED: (64, 5)}
CANDIDATES = {
    _lib.GHOST_FP64: tuple((nb, v) for nb in (32, 64) for v in (1, 2, 4)),
    # FP32 factors: the FP32 MFMA tile and the split-bf16 tile
    _lib.GHOST_MIXED: tuple((nb, v) for nb in (32, 64) for v in (1, 5)),
}
VERSION = 2   # bumped when the candidate set changes: older cache entries are not reused
_LOCK = threading.Lock()
_MEM: Dict[str, Tuple[int, int]] = {}
_cache_path = os.environ.get("GHOST_BOR_TUNE_CACHE",
                             os.path.join(os.path.expanduser("~"), ".cache", "ghost_bor_tune.json"))


def set_cache_path(path: str) -> None:
    """Use another cache file (and forget the decisions held in memory)."""
    global _cache_path
    with _LOCK:
        _cache_path = path
        _MEM.clear()


def cache_path() -> str:
    return _cache_path


def _bucket(v: int, lo: int = 6, hi: int = 12) -> int:
    """Power-of-two bucket of v, clamped to [2^lo, 2^hi]."""
    return int(2 ** max(lo, min(hi, round(math.log2(max(v, 1))))))


def key(precision: int, n: int, nsys: int, nrhs: int) -> str:
    return (f"{_lib.device_name()}|p{precision}|n{_bucket(n)}|s{_bucket(nsys, 0, 8)}"
            f"|c{_bucket(nrhs, 1, 10)}|v{VERSION}")


def _load() -> Dict[str, Tuple[int, int]]:
    try:
        with open(_cache_path) as f:
            return {k: (int(v[0]), int(v[1])) for k, v in json.load(f).items()}
    except (OSError, ValueError, TypeError, IndexError):
        return {}


def _save(cache: Dict[str, Tuple[int, int]]) -> None:
    try:
        os.makedirs(os.path.dirname(_cache_path) or ".", exist_ok=True)
        tmp = _cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, _cache_path)
    except OSError:
        pass


def measure(nb: int, variant: int, precision: int, n: int, nsys: int, nrhs: int, repeats: int = 3) -> float:
    """Best solve-phase ms over repeats of a synthetic batch: a 0.5 m sphere
    of n / 2 elements at 300 MHz, its mode systems 0..(nsys - 1) (paired, so
    about nsys systems of ~n unknowns), nrhs / 2 aspects."""
    n_el = max(8, n // 2 + 1)
    th = np.linspace(0.0, 180.0, max(1, nrhs // 2))
    t = np.linspace(0.0, math.pi, n_el + 1)
    nodes = np.column_stack([0.5 * np.sin(t), 0.5 * np.cos(t)])
    h = _lib.Handle(precision, nb=nb, residual=False)
    try:
        h.set_gemm_variant(variant)
        k = 2 * math.pi * 3e8 / 299_792_458.0
        h.add_problem(nodes, k, max(0, nsys - 1), 64, 128, 0.5, 0.5 * 376.730313668, th)
        h.build()
        h.fill()
        best_ms = float("inf")
        for _ in range(repeats):
            h.solve(0)
            h.sync()
            best_ms = min(best_ms, h.timings()["solve_ms"])
        return best_ms
    finally:
        h.close()


def best(precision: int, n: int, nsys: int, nrhs: int = 64, tune: bool = True) -> Tuple[int, int]:
    """(LU block size, GEMM tile variant) for a batch of nsys systems of up to
    n unknowns with nrhs columns: cached, else measured when tune is true,
    else the defaults."""
    k = key(precision, n, nsys, nrhs)
    with _LOCK:
        if k in _MEM:
            return _MEM[k]
        cache = _load()
        if k in cache:
            _MEM[k] = cache[k]
            return cache[k]
    if not tune:
        return DEFAULT[precision]
    shape = (_bucket(n), min(max(nsys, 1), 72), min(max(nrhs, 2), 722))
    times = {c: measure(c[0], c[1], precision, *shape) for c in CANDIDATES[precision]}
    choice = min(times, key=times.get)
    with _LOCK:
        cache = _load()
        cache[k] = choice
        _save(cache)
        _MEM[k] = choice
    return choice


def table() -> Dict[str, Tuple[int, int]]:
    """The cached decisions."""
    return dict(_load())
