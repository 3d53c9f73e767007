"""The dispatch front end end to end on the HIP path against the reference's
own outputs (fixtures from tests/golden/make_golden_dispatch.py and
make_golden_pmchwt.py): a .geo body with tapered and tabulated impedance
flags, a homogeneous dielectric body, and the adaptive frequency sweep (run
on an MI355X)."""

import os

import numpy as np
import pytest

from conftest import ROOT, pkg

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden")


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("pol", ["VV", "HH"])
def test_geo_file_with_impedance_flags_matches_reference(pol):
    bd = pkg("bor_dispatch")
    gf = pkg("geo_format")
    d = np.load(os.path.join(GOLD, "dispatch_ibc.npz"))
    snap = gf.load_geometry_snapshot(os.path.join(GOLD, "ibc_sphere.geo"))
    print(f"[case] dispatch ibc {pol}", flush=True)
    out = bd.solve_monostatic_rcs_bor(snap, [0.4, 0.7], [0.0, 45.0, 90.0, 135.0, 180.0], pol,
                                      geometry_units="meters", material_base_dir=os.path.join(GOLD, "materials"),
                                      n_modes=5, workers=1)
    assert out["metadata"]["formulation"] == str(d["formulation"])
    got_db = np.array([s["rcs_db"] for s in out["samples"]])
    got_amp = np.array([s["rcs_amp_real"] + 1j * s["rcs_amp_imag"] for s in out["samples"]])
    assert np.max(np.abs(got_db - d[f"{pol}_rcs_db"])) < 1e-7
    assert rel(got_amp, d[f"{pol}_amp"]) < 1e-9
    assert [m["modes_used"] for m in out["metadata"]["per_frequency"]] == list(d[f"{pol}_modes_used"])


def test_adaptive_frequency_sweep_matches_reference():
    bd = pkg("bor_dispatch")
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import dispatch_snapshot
    d = np.load(os.path.join(GOLD, "adaptive.npz"))
    out = bd.solve_adaptive_frequency_sweep_bor(dispatch_snapshot(), 0.2, 0.6, [0.0, 90.0], "VV",
                                                initial_points=4, max_refinements=2, db_threshold=0.5,
                                                geometry_units="meters", n_modes=4, workers=1)
    assert out["metadata"]["final_point_count"] == int(d["count"])
    assert out["metadata"]["refinement_count"] == int(d["rounds"])
    f = np.array([s["frequency_ghz"] for s in out["samples"]])
    assert np.array_equal(f, d["freqs"])
    assert np.max(np.abs(np.array([s["rcs_db"] for s in out["samples"]]) - d["rcs_db"])) < 1e-7


@pytest.mark.parametrize("pol", ["VV", "HH"])
def test_dielectric_body_matches_reference(pol):
    """A TYPE 3 sphere on one dielectric flag through the dispatch: PMCHWT
    meshed at the interior wavelength (tests/golden/make_golden_pmchwt.py)."""
    import json
    bd = pkg("bor_dispatch")
    d = np.load(os.path.join(GOLD, "dispatch_dielectric.npz"))
    snap = json.loads(str(d["snapshot"]))
    print(f"[case] dispatch dielectric {pol}", flush=True)
    out = bd.solve_monostatic_rcs_bor(snap, [0.3, 0.45], [0.0, 45.0, 90.0, 135.0, 180.0], pol,
                                      geometry_units="meters", n_modes=4, workers=1, precision="mixed")
    assert out["metadata"]["formulation"] == str(d["formulation"])
    assert out["metadata"]["precision"] == "fp64"
    got_amp = np.array([s["rcs_amp_real"] + 1j * s["rcs_amp_imag"] for s in out["samples"]])
    got_db = np.array([s["rcs_db"] for s in out["samples"]])
    assert rel(got_amp, d[f"{pol}_amp"]) < 1e-9
    assert np.max(np.abs(got_db - d[f"{pol}_rcs_db"])) < 1e-7
    assert [m["modes_used"] for m in out["metadata"]["per_frequency"]] == list(d["modes_used"])
    assert [m["n_unknowns"] for m in out["metadata"]["per_frequency"]] == list(d["n_unknowns"])
