"""Multi-region PMCHWT (PEC cores under full or partial coatings, layers and
patches, and their .geo dispatch) on the HIP path
against the reference's golden vectors
(tests/golden/make_golden_multiregion.py), the FP64 oracle and the layered
Mie series (run on an MI355X)."""

import math
import os

import numpy as np
import pytest

from conftest import pkg

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_cross_operators_match_reference():
    d = load("coated1_multiregion")
    bs = pkg("bor_solver")
    nm = int(d["m_max"])
    med = (complex(d["eps"]), complex(d["mu"]))
    so = bs.BorPecSolver(d["outer"], float(d["freq_hz"]), medium=med)
    sc = bs.BorPecSolver(d["core"], float(d["freq_hz"]), medium=med)
    print("[case] cross operators coated1", flush=True)
    for key, X in (("oc", bs.BorCrossOperators(so, sc)), ("co", bs.BorCrossOperators(sc, so))):
        for m in range(-nm, nm + 1):
            assert rel(X.assemble_T(m, nm), d[f"T{key}_{m}"]) < 1e-11, (key, m, rel(X.assemble_T(m, nm), d[f"T{key}_{m}"]))
            assert rel(X.assemble_P(m, nm), d[f"P{key}_{m}"]) < 1e-11, (key, m, rel(X.assemble_P(m, nm), d[f"P{key}_{m}"]))


@pytest.mark.parametrize("pairing", [True, False])
def test_region_systems_currents_farfield_match_reference(pairing):
    """Reduced systems and right-hand sides as assembled on the device, then
    the currents and each mode's far field (one coating layer)."""
    d = load("coated1_multiregion")
    bs = pkg("bor_solver")
    L = pkg("_lib")
    nm = int(d["m_max"])
    surf, reg = bs.coated_layers([d["outer"]], d["core"], [complex(d["eps"])], [complex(d["mu"])])
    S = bs.MultiRegionBor(surf, reg, float(d["freq_hz"]))
    print(f"[case] coated1 region set pairing={pairing}", flush=True)
    h = L.Handle(L.GHOST_FP64, pairing=pairing)
    try:
        rs = S.build(h, d["thetas"], nm)
        h.build()
        h.fill()
        h.sync()
        for m in range(-nm, nm + 1):
            A, B = h.region_system(rs, m)
            assert A.shape == d[f"A_{m}"].shape, m
            assert rel(A, d[f"A_{m}"]) < 1e-11, (m, rel(A, d[f"A_{m}"]))
            assert rel(B, d[f"B_{m}"]) < 1e-13, (m, rel(B, d[f"B_{m}"]))
        h.solve(0)
        h.farfield()
        h.sync()
        assert np.max(h.residuals()) < 1e-12
        dF = h.region_dF(rs)
        for m in range(-nm, nm + 1):
            X = h.region_solution(rs, m)
            assert rel(X, d[f"X_{m}"]) < 1e-9, (m, rel(X, d[f"X_{m}"]))
            assert rel(dF[m + nm], d[f"dF_{m}"]) < 1e-9, (m, rel(dF[m + nm], d[f"dF_{m}"]))
    finally:
        h.close()


def test_coated_entry_points_match_reference():
    d = load("coated1_multiregion")
    bs = pkg("bor_solver")
    nm, f, th = int(d["m_max"]), float(d["freq_hz"]), d["thetas"]
    eps, mu = complex(d["eps"]), complex(d["mu"])
    print("[case] solve_bor_coated_pec / coated_n", flush=True)
    a = bs.solve_bor_coated_pec(d["outer"], d["core"], f, th, eps, mu, n_modes=nm)
    b = bs.solve_bor_coated_n_pec([d["outer"]], d["core"], f, th, [eps], [mu], n_modes=nm)
    for out, pre in ((a, "coated_"), (b, "coated_n_")):
        for key in ("amp_vv", "amp_hh"):
            assert rel(np.array(out[key]), d[pre + key]) < 1e-9, (pre, key)
        assert out["modes_used"] == int(d["coated_modes_used"])
        assert out["n_unknowns"] == int(d["coated_n_unknowns"])
        assert out["linear_residual"] < 1e-12
    assert a["formulation"] == "pmchwt-coated" and b["formulation"] == "pmchwt-coated-1layer"


def test_two_layers_match_reference():
    d = load("coated2_multiregion")
    bs = pkg("bor_solver")
    print("[case] solve_bor_coated2_pec", flush=True)
    out = bs.solve_bor_coated2_pec(d["outer"], d["mid"], d["core"], float(d["freq_hz"]), d["thetas"],
                                   complex(d["eps_inner"]), complex(d["mu_inner"]), complex(d["eps_outer"]),
                                   complex(d["mu_outer"]), n_modes=int(d["m_max"]))
    for key in ("amp_vv", "amp_hh"):
        assert rel(np.array(out[key]), d[key]) < 1e-9, key
    assert out["modes_used"] == int(d["modes_used"]) and out["n_unknowns"] == int(d["n_unknowns"])


def test_coated_sphere_against_oracle_and_mie():
    """PEC sphere a = 0.3 m under a lossy magnetodielectric layer to 0.4 m,
    300 MHz, 24 / 18 segments: the device result equals the FP64 oracle and
    lies within the discretisation error of the layered Mie series."""
    import bor_oracle as bo
    import mie
    bs = pkg("bor_solver")
    eps, mu = 2.5 - 0.2j, 1.2 - 0.05j
    outer, core = bs.sphere_generatrix(0.4, 24), bs.sphere_generatrix(0.3, 18)
    th = [0.0, 60.0, 120.0, 180.0]
    print("[case] coated sphere vs oracle and Mie", flush=True)
    out = bs.solve_bor_coated_pec(outer, core, 300e6, th, eps, mu, n_modes=6, mode_tol=0.0)
    surf, reg = bo.coated_layers([outer], core, [eps], [mu])
    ref = bo.solve_multiregion(bo.MultiRegion(surf, reg, 300e6, 6), th, 0.0)
    for key in ("amp_vv", "amp_hh"):
        assert rel(np.array(out[key]), np.array(ref[key])) < 1e-9, key
    mie_db = 10.0 * math.log10(mie.sigma_coated_pec_sphere(0.3, 0.4, eps, mu, 300e6))
    for key in ("sigma_vv", "sigma_hh"):
        assert np.max(np.abs(10.0 * np.log10(np.array(out[key])) - mie_db)) < 0.1, key


def test_partial_coating_and_patch_match_reference():
    """A coated cap (one junction), a coated band (two junctions) and a
    second layer over part of a coating, end to end against the reference's
    solve_bor_partial_coating / solve_bor_coating_patch
    (tests/golden/make_golden_multiregion_dispatch.py)."""
    bs = pkg("bor_solver")
    d = load("partial_multiregion")
    nm, f, th = int(d["m_max"]), float(d["freq_hz"]), d["thetas"]
    for tag, bares in (("cap", [d["cap_bare"]]), ("band", [d["band_bare0"], d["band_bare1"]])):
        print(f"[case] solve_bor_partial_coating {tag}", flush=True)
        out = bs.solve_bor_partial_coating(d[tag + "_iface"], d[tag + "_covered"], bares, f, th, complex(d["eps"]),
                                           complex(d["mu"]), n_modes=nm)
        for key in ("amp_vv", "amp_hh"):
            assert rel(np.array(out[key]), d[tag + "_" + key]) < 1e-9, (tag, key)
        assert out["modes_used"] == int(d[tag + "_modes_used"])
        assert out["n_unknowns"] == int(d[tag + "_n_unknowns"])
        assert out["n_junctions"] == int(d[tag + "_n_junctions"])
        assert out["linear_residual"] < 1e-12
    with pytest.raises(NotImplementedError):
        bs.solve_bor_partial_coating(d["cap_iface"], d["cap_covered"], [d["cap_bare"]], f, th, 2.0, 1.0,
                                     bare_zs=[50.0], n_modes=nm)
    d = load("patch_multiregion")
    print("[case] solve_bor_coating_patch", flush=True)
    out = bs.solve_bor_coating_patch(d["patch"], d["mid_cov"], d["mid_bare"], d["core"], float(d["freq_hz"]),
                                     d["thetas"], complex(d["eps_inner"]), complex(d["mu_inner"]),
                                     complex(d["eps_patch"]), complex(d["mu_patch"]), n_modes=int(d["m_max"]))
    for key in ("amp_vv", "amp_hh"):
        assert rel(np.array(out[key]), d[key]) < 1e-9, key
    assert out["n_unknowns"] == int(d["n_unknowns"]) and out["n_junctions"] == int(d["n_junctions"])


@pytest.mark.parametrize("name", ["coated", "partial", "layered", "patch", "layered_n"])
def test_dispatch_multi_region_layouts_match_reference(name):
    """.geo snapshots of the multi-region layouts through
    solve_monostatic_rcs_bor at two frequencies against the reference's
    dispatch (samples, modes used, unknowns, formulation label)."""
    import sys
    sys.path.insert(0, GOLD)
    from make_golden_multiregion_dispatch import snapshots, THETAS, NM
    bd = pkg("bor_dispatch")
    d = load("dispatch_multiregion")
    print(f"[case] dispatch {name}", flush=True)
    res = bd.solve_monostatic_rcs_bor(snapshots()[name], [0.3, 0.36], list(THETAS), "VV", geometry_units="meters",
                                      n_modes=NM, workers=1)
    amp = np.array([s["rcs_amp_real"] + 1j * s["rcs_amp_imag"] for s in res["samples"]])
    assert rel(amp, d[name + "_amp"]) < 1e-9
    assert np.max(np.abs(np.array([s["rcs_db"] for s in res["samples"]]) - d[name + "_db"])) < 1e-7
    meta = res["metadata"]
    assert [m["modes_used"] for m in meta["per_frequency"]] == list(d[name + "_modes_used"])
    assert [m["n_unknowns"] for m in meta["per_frequency"]] == list(d[name + "_n_unknowns"])
    assert meta["formulation"] == str(d[name + "_formulation"])
    assert meta["geometry_kind"] == str(d[name + "_kind"])
