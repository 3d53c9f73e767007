"""Autotuner for the batched LU: block size and GEMM tile.

The LU block size ``nb`` sets the panel width (latency-bound work, grows with
nb) against the K depth of the MFMA trailing updates (arithmetic intensity
nb/4 flop/byte on the HBM-resident trailing matrix, grows with nb); the GEMM
tile variant trades LDS reuse against occupancy and edge waste on small
systems.  The best pair depends on the system size n, the batch size and the
precision, so it is measured: for a size bucket the tuner builds a synthetic
batch of sphere systems of about that size, times the solve phase (HIP
events on the handle's stream) for each candidate, and caches the winner
per (device, precision, n bucket, batch bucket) in a JSON file.
"""

from __future__ import annotations

import json
import math
import os
import threading
from typing import Dict, Optional, Tuple

import numpy as np

from . import _lib

CANDIDATES = tuple((nb, v) for nb in (32, 64) for v in (0, 1, 2))
_CACHE_PATH = os.environ.get("GHOST_BOR_TUNE_CACHE",
                             os.path.join(os.path.expanduser("~"), ".cache", "ghost_bor_tune.json"))
_LOCK = threading.Lock()
_MEM: Dict[str, Tuple[int, int]] = {}


def _bucket(n: int) -> int:
    """Power-of-two size bucket (64, 128, ..., 4096)."""
    return int(2 ** max(6, min(12, round(math.log2(max(n, 1))))))


def _device_tag() -> str:
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.get_device_name(0).replace(" ", "_")
    except Exception:
        pass
    return "hip-device"


def _key(precision: int, n: int, nsys: int) -> str:
    return f"{_device_tag()}|p{precision}|n{_bucket(n)}|s{_bucket(nsys)}"


def _load_cache() -> Dict[str, Tuple[int, int]]:
    try:
        with open(_CACHE_PATH) as f:
            return {k: (int(v[0]), int(v[1])) for k, v in json.load(f).items()}
    except (OSError, ValueError, TypeError, IndexError):
        return {}


def _save_cache(cache: Dict[str, Tuple[int, int]]) -> None:
    try:
        os.makedirs(os.path.dirname(_CACHE_PATH), exist_ok=True)
        tmp = _CACHE_PATH + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, _CACHE_PATH)
    except OSError:
        pass


def measure(nb: int, variant: int, precision: int, n: int, nsys: int, nrhs: int, repeats: int = 2) -> float:
    """Solve-phase ms of a synthetic batch (sphere systems of ~n unknowns)."""
    n_el = max(8, n // 2 + 1)
    m_max = max(0, (nsys - 1) // 2)
    th = np.linspace(0.0, 180.0, max(1, nrhs // 2))
    nodes = np.column_stack([0.5 * np.sin(np.linspace(0, math.pi, n_el + 1)),
                             0.5 * np.cos(np.linspace(0, math.pi, n_el + 1))])
    h = _lib.Handle(precision, nb=nb, residual=False)
    h.set_gemm_variant(variant)
    try:
        k = 2 * math.pi * 3e8 / 299_792_458.0
        h.add_problem(nodes, k, m_max, 64, 128, 0.5, 0.5 * 376.730313668, th)
        h.build()
        best = float("inf")
        for _ in range(repeats):
            h.fill()
            h.solve(1 if precision == _lib.GHOST_MIXED else 0)
            h.sync()
            best = min(best, h.timings()["solve_ms"])
        return best
    finally:
        h.close()


def best(precision: int, n: int, nsys: int, nrhs: int = 64, tune: bool = True) -> Tuple[int, int]:
    """Cached best (LU block size, GEMM tile variant) for this shape; measured
    on first use when ``tune`` is true, else the defaults (64, 1)."""
    key = _key(precision, n, nsys)
    with _LOCK:
        if key in _MEM:
            return _MEM[key]
        cache = _load_cache()
        if key in cache:
            _MEM[key] = cache[key]
            return cache[key]
    if not tune:
        return (64, 1)
    times = {c: measure(c[0], c[1], precision, _bucket(n), min(nsys, 33), min(nrhs, 128)) for c in CANDIDATES}
    choice = min(times, key=times.get)
    with _LOCK:
        cache = _load_cache()
        cache[key] = choice
        _save_cache(cache)
        _MEM[key] = choice
    return choice


def table() -> Dict[str, Tuple[int, int]]:
    """The cached decisions."""
    return dict(_load_cache())
