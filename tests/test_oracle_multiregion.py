"""The oracle's multi-region PMCHWT (PEC cores under coating layers),
pinned against the reference's own vectors
(tests/golden/make_golden_multiregion.py) and the layered-sphere Mie series
(an admittance recursion, not the integral equation)."""

import math
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
from oracle import bor_oracle as bo  # noqa: E402
from oracle import mie  # noqa: E402

GOLD = os.path.join(ROOT, "tests", "golden")


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def coated1():
    d = np.load(os.path.join(GOLD, "coated1_multiregion.npz"))
    surf, reg = bo.coated_layers([d["outer"]], d["core"], [complex(d["eps"])], [complex(d["mu"])])
    return d, bo.MultiRegion(surf, reg, float(d["freq_hz"]), int(d["m_max"]))


def test_cross_operators_match_reference(coated1):
    d, S = coated1
    nm = int(d["m_max"])
    for key, X in (("oc", S.X[(1, 0, 1)]), ("co", S.X[(1, 1, 0)])):
        ref = {(int(e), int(f)) for e, f, _ in d[f"near_{key}"]}
        assert set(X.kinds) == ref and all(v is None for v in X.kinds.values())
        for m in range(-nm, nm + 1):
            assert _rel(X.T(m), d[f"T{key}_{m}"]) < 1e-12, (key, m)
            assert _rel(X.P(m), d[f"P{key}_{m}"]) < 1e-12, (key, m)


def test_region_system_matches_reference(coated1):
    d, S = coated1
    out = bo.solve_multiregion(S, d["thetas"], 1e-6, keep=True)
    nm = int(d["m_max"])
    for m in range(-nm, nm + 1):
        Q = S.Q(m)
        assert np.array_equal(Q, d[f"Q_{m}"]), m
        X, B, A = out["_modes"][m]
        assert _rel(A, d[f"A_{m}"]) < 1e-12, m
        assert _rel(B, d[f"B_{m}"]) < 1e-12, m
        assert _rel(X, d[f"X_{m}"]) < 1e-10, m
    for key in ("amp_vv", "amp_hh"):
        assert _rel(np.array(out[key]), d["coated_n_" + key]) < 1e-12
        # the reference's two coated entry points (an explicit 3 x 3 block
        # system and the region system) give the same far field
        assert _rel(d["coated_" + key], d["coated_n_" + key]) < 1e-12
    assert out["modes_used"] == int(d["coated_modes_used"])
    assert out["n_unknowns"] == int(d["coated_n_unknowns"])


def test_two_layers_match_reference():
    d = np.load(os.path.join(GOLD, "coated2_multiregion.npz"))
    surf, reg = bo.coated_layers([d["outer"], d["mid"]], d["core"],
                                 [complex(d["eps_inner"]), complex(d["eps_outer"])],
                                 [complex(d["mu_inner"]), complex(d["mu_outer"])])
    S = bo.MultiRegion(surf, reg, float(d["freq_hz"]), int(d["m_max"]))
    out = bo.solve_multiregion(S, d["thetas"], 1e-6, keep=True)
    for key in ("amp_vv", "amp_hh"):
        assert _rel(np.array(out[key]), d[key]) < 1e-12
    assert out["modes_used"] == int(d["modes_used"]) and out["n_unknowns"] == int(d["n_unknowns"])
    for m in (0, 1):
        assert _rel(out["_modes"][m][0], d[f"X_{m}"]) < 1e-10


def test_coated_sphere_against_mie():
    """PEC sphere a = 0.3 m under a lossy magnetodielectric layer to 0.4 m,
    300 MHz: within the discretisation error of the layered Mie series."""
    eps, mu = 2.5 - 0.2j, 1.2 - 0.05j
    surf, reg = bo.coated_layers([bo.sphere_nodes(0.4, 24)], bo.sphere_nodes(0.3, 18), [eps], [mu])
    out = bo.solve_multiregion(bo.MultiRegion(surf, reg, 300e6, 6), [0.0, 60.0, 120.0, 180.0], 0.0)
    ref = 10.0 * math.log10(mie.sigma_coated_pec_sphere(0.3, 0.4, eps, mu, 300e6))
    for key in ("sigma_vv", "sigma_hh"):
        assert np.max(np.abs(10.0 * np.log10(out[key]) - ref)) < 0.1, key


def test_layered_mie_identities():
    """A free-space layer leaves the PEC core; two layers of one medium are
    one layer; the multilayer series with one layer is the coated one."""
    f = 300e6
    assert abs(mie.sigma_coated_pec_sphere(0.3, 0.5, 1.0, 1.0, f) / mie.sigma_pec_sphere(0.3, f) - 1) < 1e-12
    e, u = 2.5 - 0.2j, 1.1 - 0.03j
    one = mie.sigma_coated_pec_sphere(0.3, 0.5, e, u, f)
    two = mie.sigma_multilayer_pec_sphere([0.3, 0.41, 0.5], [e, e], [u, u], f)
    assert abs(two / one - 1) < 1e-12
    # a vanishing layer: the PEC sphere of the outer radius
    thin = mie.sigma_coated_pec_sphere(0.5 - 1e-9, 0.5, 4.0 - 1j, 1.0, f)
    assert abs(thin / mie.sigma_pec_sphere(0.5, f) - 1) < 1e-6


def test_cross_pair_classification():
    """Touching generatrices: a shared endpoint is a corner pair with the
    shared ends named; a close pair is a tensor pair; overlap is an error."""
    a = bo.Body(np.array([[0.0, 1.0], [0.5, 0.5], [0.5, 0.0]]))
    b = bo.Body(np.array([[0.5, 0.0], [0.3, -0.4], [0.0, -0.5]]))
    kinds = bo.cross_pair_kinds(a, b)
    assert kinds[(1, 0)] == "corner10" and kinds[(0, 0)] is None
    # element 0 of a and element 1 of b are 0.92 m apart; the longer is 0.71 m
    assert (0, 1) in kinds and (0, 1) not in bo.cross_pair_kinds(a, b, near_factor=1.0)
    c = bo.Body(np.array([[0.0, 1.0], [0.5, 0.5], [0.9, 0.0]]))
    with pytest.raises(ValueError):
        bo.cross_pair_kinds(a, c)


def test_partial_coating_and_patch_layouts_match_reference():
    """The region system with the partial-coating layout (a coated cap with
    one junction, a coated band with two) and with the coating-patch layout
    gives the reference's dedicated solvers' far fields
    (tests/golden/make_golden_multiregion_dispatch.py)."""
    d = np.load(os.path.join(GOLD, "partial_multiregion.npz"))
    eps, mu = complex(d["eps"]), complex(d["mu"])
    for tag, bares in (("cap", [d["cap_bare"]]), ("band", [d["band_bare0"], d["band_bare1"]])):
        surf, reg = bo.partial_coating_layout(d[tag + "_iface"], d[tag + "_covered"], bares, eps, mu)
        out = bo.solve_multiregion(bo.MultiRegion(surf, reg, float(d["freq_hz"]), int(d["m_max"])), d["thetas"])
        for key in ("amp_vv", "amp_hh"):
            assert _rel(np.array(out[key]), d[tag + "_" + key]) < 1e-12, (tag, key)
        assert out["modes_used"] == int(d[tag + "_modes_used"])
        assert out["n_unknowns"] == int(d[tag + "_n_unknowns"])
        assert out["n_junctions"] == int(d[tag + "_n_junctions"])
    d = np.load(os.path.join(GOLD, "patch_multiregion.npz"))
    surf, reg = bo.coating_patch_layout(d["patch"], d["mid_cov"], d["mid_bare"], d["core"],
                                        complex(d["eps_inner"]), complex(d["mu_inner"]),
                                        complex(d["eps_patch"]), complex(d["mu_patch"]))
    out = bo.solve_multiregion(bo.MultiRegion(surf, reg, float(d["freq_hz"]), int(d["m_max"])), d["thetas"])
    for key in ("amp_vv", "amp_hh"):
        assert _rel(np.array(out[key]), d[key]) < 1e-12, key
    assert out["n_unknowns"] == int(d["n_unknowns"]) and out["n_junctions"] == int(d["n_junctions"])
