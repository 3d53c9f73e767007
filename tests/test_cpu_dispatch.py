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
    assert bd.layout(bd.read_pieces(coated, 1.0)) == "coated"
    bare = {"segments": [dict(s, properties=[2 + i, 2, 0, 1, 0]) for i, s in enumerate(snap["segments"])]}
    with pytest.raises(ValueError, match="partial coating"):   # TYPE 2 + 3 with no covered core
        bd.layout(bd.read_pieces(bare, 1.0))
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


def _unknowns(plan, bd, mat, fg, axis_tol):
    """Full unknown count of a planned multi-region body meshed as the
    dispatch meshes it: 2 Nn per conductor surface, 4 Nn per interface."""
    media = [mat.get_medium(fl, fg) for fl in plan.flags]
    lam = bd.C0 / (fg * 1e9)
    idx = max([1.0] + [abs(pkg("bor_solver").causal_medium(e, u)[0]) for (e, u) in media])
    nn = lambda run, l=lam / idx: len(bd.mesh_run(run, l, 50000, axis_tol)[0])
    R = plan.runs
    if plan.kind == "coated":
        return 4 * nn(R["interface"][0]) + 2 * nn(R["core"][0])
    if plan.kind == "partial":
        return 4 * nn(R["interface"][0]) + 2 * nn(R["core"][0]) + sum(2 * nn(r, lam) for r in R["bare"])
    if plan.kind == "layered":
        return (4 * nn(R["patch"][0]) + 4 * nn(R["mid"][0]) + sum(4 * nn(r) for r in R["exposed"])
                + 2 * nn(R["core"][0]))
    return sum(4 * nn(r) for r in R["interfaces"]) + 2 * nn(R["core"][0])


def test_multi_region_layouts_and_meshes_match_reference():
    """The coated, partially coated, two-layer, patch and three-layer
    snapshots classify as the reference classifies them and mesh to the
    reference's unknown counts at both frequencies
    (tests/golden/make_golden_multiregion_dispatch.py)."""
    from make_golden_multiregion_dispatch import snapshots
    bd = pkg("bor_dispatch")
    mats = pkg("materials")
    d = np.load(os.path.join(ROOT, "tests", "golden", "dispatch_multiregion.npz"))
    for name, snap in snapshots().items():
        pieces = bd.read_pieces(snap, 1.0)
        kind = bd.layout(pieces)
        assert kind == str(d[name + "_kind"]), name
        diag = max(float(np.ptp(np.vstack([c.pts for c in pieces]), axis=0).max()), 1e-9)
        plan = bd.region_plan(kind, pieces, max(1e-12, 1e-9 * diag), 1e-6 * diag)
        lib = mats.MaterialLibrary.from_entries([], snap["dielectrics"], base_dir=ROOT)
        got = [_unknowns(plan, bd, lib, fg, 1e-6 * diag) for fg in (0.3, 0.36)]
        assert got == list(d[name + "_n_unknowns"]), (name, got)


def test_multi_region_validation_errors():
    from make_golden_multiregion_dispatch import snapshots
    bd = pkg("bor_dispatch")
    snaps = snapshots()
    bad = dict(snaps["coated"])
    bad["segments"] = [dict(bad["segments"][0]), dict(bad["segments"][1], properties=[4, 1, 0, 2, 0])]
    with pytest.raises(ValueError):          # interface and core on different coating flags
        bd.layout(bd.read_pieces(bad, 1.0))
    bad = dict(snaps["layered"])
    bad["segments"] = [bad["segments"][0], dict(bad["segments"][1], properties=[5, 1, 0, 2, 2]),
                       bad["segments"][2]]
    with pytest.raises(ValueError):          # a TYPE 5 pair needs two different flags
        bd.layout(bd.read_pieces(bad, 1.0))
    bad = dict(snaps["layered_n"])
    bad["segments"] = [bad["segments"][0], dict(bad["segments"][1], properties=[5, 1, 0, 3, 4]),
                       bad["segments"][2], bad["segments"][3]]
    pieces = bd.read_pieces(bad, 1.0)
    with pytest.raises(ValueError):          # the layer flags do not lead to the core's
        bd.region_plan(bd.layout(pieces), pieces, 1e-9, 1e-6)
