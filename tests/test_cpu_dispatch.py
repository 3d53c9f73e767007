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
    assert bd.layout(bd.read_pieces(diel, 1.0)) == "dielectric"
    two = {"segments": [dict(s, properties=[3, 2, 0, 1 + i, 0]) for i, s in enumerate(snap["segments"])]}
    with pytest.raises(ValueError, match="pos_mat"):
        bd.layout(bd.read_pieces(two, 1.0))
    coated = {"segments": [dict(s, properties=[3 + i, 2, 0, 1, 0]) for i, s in enumerate(snap["segments"])]}
    with pytest.raises(NotImplementedError):
        bd.layout(bd.read_pieces(coated, 1.0))
    with pytest.raises(ValueError, match="TYPE"):
        bd.layout(bd.read_pieces({"segments": [dict(s, properties=[7, 2, 0, 0, 0]) for s in snap["segments"]]},
                                 1.0))
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


def test_dielectric_body_meshed_at_the_interior_wavelength():
    """TYPE 3 sphere on one dielectric flag: the element counts the
    reference meshed at lambda0 / |m| (tests/golden/dispatch_dielectric.npz
    holds its 4 Nn unknowns per frequency)."""
    import json
    bd = pkg("bor_dispatch")
    bs = pkg("bor_solver")
    mt = pkg("materials")
    d = np.load(os.path.join(ROOT, "tests", "golden", "dispatch_dielectric.npz"))
    snap = json.loads(str(d["snapshot"]))
    pieces = bd.read_pieces(snap, 1.0)
    assert bd.layout(pieces) == "dielectric"
    diag = max(float(np.ptp(np.vstack([c.pts for c in pieces]), axis=0).max()), 1e-9)
    run = bd.single_run(pieces, "TYPE 3", max(1e-12, 1e-9 * diag))
    lib = mt.MaterialLibrary.from_entries([], snap["dielectrics"], ".")
    for fg, nu in zip((0.3, 0.45), d["n_unknowns"]):
        eps, mu = lib.get_medium(run[0].pos, fg)
        m_idx, _ = bs.causal_medium(eps, mu)
        pts, _, _ = bd.mesh_run(run, bd.C0 / (fg * 1e9) / max(1.0, abs(m_idx)), 50000, 1e-6 * diag)
        assert 4 * len(pts) == int(nu)


def test_penetrable_job_validation():
    bs = pkg("bor_solver")
    nodes = bs.sphere_generatrix(0.3, 8)
    job = {"points": nodes, "freq_hz": 3e8, "thetas_deg": [0.0], "medium": (2.0, 1.0)}
    with pytest.raises(ValueError, match="FP64"):
        bs.solve_bor_batch([job], precision="mixed")
    with pytest.raises(ValueError, match="impedance"):
        bs.solve_bor_batch([dict(job, zs=50.0)])
    m, er = bs.causal_medium(3.0 - 0.4j, 1.3 - 0.05j)
    assert m.imag <= 0.0 and m.real > 0.0 and abs(m * m - (3.0 - 0.4j) * (1.3 - 0.05j)) < 1e-14
    assert abs(er - (1.3 - 0.05j) / m) < 1e-15
