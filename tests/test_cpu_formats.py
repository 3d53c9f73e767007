"""Host-side front-end pieces against the reference's own outputs (fixtures
written by tests/golden/make_golden_dispatch.py): the .geo reader and
writer, the material library, the radar-frame az/el grid and the .grim
export.  None of these needs a GPU."""

import json
import os

import numpy as np
import pytest

from conftest import ROOT, pkg

GOLD = os.path.join(ROOT, "tests", "golden")


def test_geo_parse_matches_reference():
    gf = pkg("geo_format")
    with open(os.path.join(GOLD, "ibc_sphere.geo")) as f:
        text = f.read()
    snap = gf.build_geometry_snapshot(*gf.parse_geometry(text))
    with open(os.path.join(GOLD, "geo_parse.json")) as f:
        ref = json.load(f)
    assert snap == ref
    assert snap == gf.load_geometry_snapshot(os.path.join(GOLD, "ibc_sphere.geo"))


def test_geo_text_round_trip():
    gf = pkg("geo_format")
    with open(os.path.join(GOLD, "ibc_sphere.geo")) as f:
        parsed = gf.parse_geometry(f.read())
    again = gf.parse_geometry(gf.build_geometry_text(*parsed))
    assert gf.build_geometry_snapshot(*again) == gf.build_geometry_snapshot(*parsed)


@pytest.mark.parametrize("text, what", [
    ("Title: t\n0 0 1 1\n", "before any"),
    ("Segment: a 2\nproperties: 2 0 0\n", "5"),
    ("Segment: a 2\nproperties: 2 0 0 0 0\n0 1 2\n", "4 numbers"),
    ("Segment: a 2\nIBCS:\n3 wobbly 1 2 3 4\n", "kind"),
    ("Segment: a 2\nIBCS:\n57 1\n", "flag alone"),
    ("Segment: a 2\nIBCS:\nx linear 1 2 3 4\n", "integer flag"),
])
def test_geo_parse_errors(text, what):
    gf = pkg("geo_format")
    with pytest.raises(ValueError, match=what):
        gf.parse_geometry(text)


def test_material_library_matches_reference():
    mt = pkg("materials")
    d = np.load(os.path.join(GOLD, "materials.npz"))
    lib = mt.MaterialLibrary.from_entries(
        [["3", "linear", "50", "-10", "300", "-80"], ["4", "exp", "20", "0", "400", "-50"],
         ["5", "constant", "77", "-7", "0", "0"], ["6", "cosine", "0", "0", "200", "100"], ["57"]],
        [["1", "2.5", "-0.1", "1.0", "0.0"], ["2", "0", "0"], ["61"]],
        base_dir=os.path.join(GOLD, "materials"))
    got = np.array([lib.get_impedance(int(fl), f, arc_s=s) for fl, f, s in d["imp_args"]])
    assert np.max(np.abs(got - d["imp"])) <= 1e-12 * np.max(np.abs(d["imp"]))
    med = np.array([lib.get_medium(fl, f) for fl in (1, 2, 61, 0, 9) for f in (0.05, 0.4, 1.0, 2.0)])
    assert np.max(np.abs(med - d["med"])) <= 1e-14
    # the same out-of-range lookups are reported, once each (the wording is this build's own)
    assert len(lib.warnings) == len(d["warnings"])
    for mine, ref in zip(lib.warnings, d["warnings"]):
        assert mine.split()[2] == str(ref).split()[2] and mine.split()[5] == str(ref).split()[5]
    with pytest.raises(ValueError):
        mt.MaterialLibrary.from_entries([["3", "linear", "1"]], [], ".")


def _result_from_rows(rows, pol):
    return {"polarization": pol, "samples": [
        {"frequency_ghz": r[0], "theta_inc_deg": r[1], "rcs_amp_real": r[2], "rcs_amp_imag": r[3]} for r in rows]}


def test_az_el_grid_matches_reference():
    bd = pkg("bor_dispatch")
    d = np.load(os.path.join(GOLD, "az_el.npz"))
    g = bd.bor_az_el_grid(_result_from_rows(d["vv"], "VV"), _result_from_rows(d["hh"], "HH"),
                          np.arange(-60.0, 61.0, 20.0), np.arange(-30.0, 31.0, 15.0),
                          axis_az_deg=10.0, axis_el_deg=20.0)
    assert np.max(np.abs(g["theta_map_deg"] - d["theta_map"])) < 1e-10
    for ch in ("VV", "HH", "VH"):
        assert np.max(np.abs(g["amp"][ch] - d[ch])) <= 1e-12 * np.max(np.abs(d["VV"])), ch
    with pytest.raises(ValueError):
        bd.bor_az_el_grid(_result_from_rows(d["hh"], "HH"), _result_from_rows(d["vv"], "VV"), [0.0], [0.0])


def test_az_el_grid_reduces_to_the_sweep_on_a_vertical_axis():
    """Axis straight up: every look's meridian basis is the radar basis, so
    VV = F_VV(theta), HH = F_HH(theta), VH = 0 with theta = 90 - el."""
    bd = pkg("bor_dispatch")
    d = np.load(os.path.join(GOLD, "az_el.npz"))
    el = np.array([-60.0, -20.0, 0.0, 30.0, 70.0])
    g = bd.bor_az_el_grid(_result_from_rows(d["vv"], "VV"), _result_from_rows(d["hh"], "HH"),
                          [15.0, 100.0], el, axis_az_deg=0.0, axis_el_deg=90.0)
    vv = d["vv"][d["vv"][:, 0] == d["vv"][0, 0]]
    th = 90.0 - el
    Fv = np.interp(th, vv[:, 1], vv[:, 2]) + 1j * np.interp(th, vv[:, 1], vv[:, 3])
    assert np.allclose(g["amp"]["VV"][0, :, 0], Fv, rtol=1e-12, atol=0)
    assert np.max(np.abs(g["amp"]["VH"])) < 1e-12 * np.max(np.abs(Fv))


def test_grim_export_matches_reference(tmp_path):
    gr = pkg("grim_format")
    with open(os.path.join(GOLD, "pec_sweep_vv.json")) as f:
        res = json.load(f)
    path = gr.export_result_to_grim(res, str(tmp_path / "sweep"), source_path="s", history="h")[0]
    assert path.endswith(".grim")
    ref = np.load(os.path.join(GOLD, "grim_mono.npz"))
    with np.load(path) as z:
        assert sorted(z.files) == sorted(ref.files)
        for k in ref.files:
            a, b = z[k], ref[k]
            if a.dtype.kind == "f":
                assert np.array_equal(a, b, equal_nan=True), k
            else:
                assert np.array_equal(a, b), k
