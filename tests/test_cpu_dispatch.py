"""Host-side dispatch logic (snapshot -> chains -> generatrix) against the
reference's own meshing (golden fixture) and its validation errors."""

import math
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, pkg

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from make_golden import dispatch_snapshot  # noqa: E402


def test_mesh_matches_reference():
    bd = pkg("bor_dispatch")
    d = np.load(os.path.join(ROOT, "tests", "golden", "dispatch_sphere.npz"))
    snap = dispatch_snapshot()
    pieces = bd.read_pieces(snap, 1.0)
    assert bd.layout(pieces) == "conductor"
    diag = max(float(np.ptp(np.vstack([c.pts for c in pieces]), axis=0).max()), 1e-9)
    run = bd.single_run(pieces, "TYPE 2", max(1e-12, 1e-9 * diag))
    bd.check_generatrix(run, "body", 1e-6 * diag)
    pts, owner, arc = bd.mesh_run(run, bd.C0 / 0.3e9, 50000, 1e-6 * diag)
    assert len(owner) == len(pts) - 1 and np.all((arc > 0) & (arc < 1))
    assert pts.shape == d["nodes"].shape
    assert np.max(np.abs(pts - d["nodes"])) < 1e-15


def test_validation_errors():
    bd = pkg("bor_dispatch")
    snap = dispatch_snapshot()
    with pytest.raises(ValueError):
        bd.solve_monostatic_rcs_bor(snap, [], [0.0], "VV")
    with pytest.raises(ValueError):
        bd.solve_monostatic_rcs_bor(snap, [0.3], [200.0], "VV")
    with pytest.raises(ValueError):
        bd.solve_monostatic_rcs_bor(snap, [0.3], [0.0], "XX")
    rev = {"segments": [dict(s, point_pairs=[{"x1": p["x2"], "y1": p["y2"], "x2": p["x1"], "y2": p["y1"]}
                                              for p in reversed(s["point_pairs"])])
                        for s in reversed(snap["segments"])]}
    run = bd.single_run(bd.read_pieces(rev, 1.0), "TYPE 2", 1e-9)
    with pytest.raises(ValueError, match="traversed"):
        bd.check_generatrix(run, "body", 1e-6)
    diel = {"segments": [dict(s, properties=[3, 2, 0, 1, 0]) for s in snap["segments"]]}
    with pytest.raises(NotImplementedError):
        bd.layout(bd.read_pieces(diel, 1.0))
    abort = __import__("threading").Event()
    abort.set()
    with pytest.raises(InterruptedError):
        bd.solve_monostatic_rcs_bor(snap, [0.3], [0.0], "VV", geometry_units="meters", abort_event=abort)


def test_element_count_rule():
    bd = pkg("bor_dispatch")
    assert bd.element_count(5, 1.0, 0.1) == 5
    assert bd.element_count(-10, 1.0, 0.1) == 100
    assert bd.element_count(0, 1.0, 0.1) == 200
    assert bd.element_count(0, 1.0, 1e-9) == 2000
