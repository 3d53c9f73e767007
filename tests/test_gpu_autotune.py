"""The LU autotuner in the product path (run on an MI355X)."""

import json

import numpy as np
import pytest

import bor_oracle as bo
from conftest import pkg

pytestmark = pytest.mark.gpu


def test_tuned_tiles_are_cached_applied_and_change_nothing(tmp_path, monkeypatch):
    at = pkg("autotune")
    bs = pkg("bor_solver")
    L = pkg("_lib")
    orig = at.cache_path()
    at.set_cache_path(str(tmp_path / "tune.json"))
    try:
        name = L.device_name()
        print(f"[case] device {name}", flush=True)
        assert "gfx" in name
        nodes = bo.sphere_nodes(0.5, 60)
        job = {"points": nodes, "freq_hz": 3e8, "thetas_deg": np.linspace(0, 180, 31), "n_modes": 9}
        out = bs.solve_bor_batch([job], mode_tol=0.0)[0]
        tiles = tuple(out["lu_tiles"])
        assert tiles in at.CANDIDATES[L.GHOST_FP64]
        with open(tmp_path / "tune.json") as f:
            cache = json.load(f)
        assert list(cache.values()) == [list(tiles)]
        key = next(iter(cache))
        assert key.startswith(name) and "|p0|" in key
        # a second call reads the cache: no measurement
        calls = []
        monkeypatch.setattr(at, "measure", lambda *a, **k: calls.append(a) or 1.0)
        again = bs.solve_bor_batch([job], mode_tol=0.0)[0]
        assert tuple(again["lu_tiles"]) == tiles and not calls
        assert np.array_equal(np.array(again["amp_vv"]), np.array(out["amp_vv"]))
        # the tuned tile gives the default tile's result (FP64 rounding only)
        ref = bs.solve_bor_batch([job], mode_tol=0.0, nb=64)[0]
        rel = np.max(np.abs(np.array(ref["amp_vv"]) - np.array(out["amp_vv"]))) / np.max(np.abs(ref["amp_vv"]))
        assert rel < 1e-12
    finally:
        at.set_cache_path(orig)
